// Common HIP helpers for the laminocam gfx950 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <new>
#include <string>

#include "laminocam.h"

namespace lc {

// complex storage types
template <typename T> struct cpx_t;
template <> struct cpx_t<float> { using type = float2; };
template <> struct cpx_t<double> { using type = double2; };
template <typename T> using cpx = typename cpx_t<T>::type;

template <typename C> __host__ __device__ __forceinline__ C cmake(decltype(C{}.x) a, decltype(C{}.x) b) {
    C z;
    z.x = a;
    z.y = b;
    return z;
}

// four-wide vector of the element type (block copies of two complex values)
template <typename T> struct vec4_t;
template <> struct vec4_t<float> { using type = float4; };
template <> struct vec4_t<double> { using type = double4; };
template <typename T> using vec4 = typename vec4_t<T>::type;
template <typename V>
__host__ __device__ __forceinline__ V vmake4(decltype(V{}.x) a, decltype(V{}.x) b,
                                             decltype(V{}.x) c, decltype(V{}.x) d) {
    V v;
    v.x = a;
    v.y = b;
    v.z = c;
    v.w = d;
    return v;
}
// streaming (nontemporal) four-wide load and store, for vectors read or
// written once per pass: the data sets are far larger than the caches, and
// the hint measures about 3% faster on a read-4 / write-2 stream
template <typename T> using native4 = T __attribute__((ext_vector_type(4)));
template <typename V> __device__ __forceinline__ V ld_stream(const V *p) {
    using T = decltype(V{}.x);
    const native4<T> t = __builtin_nontemporal_load(reinterpret_cast<const native4<T> *>(p));
    return vmake4<V>(t.x, t.y, t.z, t.w);
}
template <typename V> __device__ __forceinline__ void st_stream(V *p, const V &v) {
    using T = decltype(V{}.x);
    const native4<T> t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<native4<T> *>(p));
}

__host__ __device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ double2 operator+(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 operator-(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// error plumbing: library code throws lc::Error; the C ABI converts to status
struct Error : public std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define LC_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t _e = (call);                                                       \
        if (_e != hipSuccess)                                                         \
            throw ::lc::Error(LC_EHIP, std::string(#call) + ": " +                    \
                                           hipGetErrorString(_e) + " at " + __FILE__ + \
                                           ":" + std::to_string(__LINE__));            \
    } while (0)

#define LC_REQUIRE(cond, msg)                                                         \
    do {                                                                              \
        if (!(cond)) throw ::lc::Error(LC_EINVAL, std::string(msg));                 \
    } while (0)

// the calling thread's message for lc_last_error
std::string &last_error();

// runs f at the C boundary: an exception becomes a status code and a message
template <typename F> int c_guard(F &&f) {
    try {
        f();
        last_error().clear();
        return LC_OK;
    } catch (const Error &e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        last_error() = "host allocation failed";
        return LC_ENOMEM;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return LC_EINTERNAL;
    } catch (...) {
        last_error() = "unknown error";
        return LC_EINTERNAL;
    }
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// orders a wave's LDS writes before its later LDS reads of other lanes' data
// (and reads before later overwrites); a wave executes in lockstep, so no
// barrier across waves is needed
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// blockIdx remap so that consecutive logical blocks land on the same XCD.
// Hardware dispatch sends block b to XCD (b mod 8); giving XCD x a contiguous
// run of logical ids lets its private L2 share the cache lines neighbouring
// blocks touch.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    constexpr int NX = 8;
    if (nb < NX) return b;
    const int x = b % NX, q = nb / NX, rem = nb % NX;
    const int before = x * q + (x < rem ? x : rem);
    return before + b / NX;
}

} // namespace lc

// LDS-resident mixed-radix Stockham FFT core for gfx950.
//
// A workgroup transforms `nl` lines of length L held in LDS.  Two layouts:
//   ROW (LFAST=false): element e of line l at swz(l*ls + e)   (ls >= L)
//   COL (LFAST=true):  element e of line l at swz(e*nl + l)   (lines interleaved)
// swz() permutes addresses inside aligned 32-element groups so the strided
// Stockham writes spread over the 64 LDS banks (LDS buffers are allocated in
// multiples of 32 elements).  Each stage of radix R reads its R inputs into
// registers, synchronises, then writes the outputs in autosorted (Stockham)
// order back into the same buffer, so one LDS buffer suffices.
//
// Stage with radix R after Ns = product of the previous radices:
//   for j < L/R, k = j mod Ns:
//     v_r = in[j + r L/R] * w_L^{k r L/(Ns R)}      (w_L = e^{-2 pi i/L})
//     V   = DFT_R(v)
//     out[(j/Ns) Ns R + k + r Ns] = V_r
// Twiddles come from a table tw[k] = e^{-2 pi i k / L} generated in FP64.
#pragma once

#include "lc_common.hpp"

namespace lc {

constexpr int FFT_MAX_STAGES = 16;

struct FftPlan {
    int L;
    int nst;
    int rad[FFT_MAX_STAGES];
};

// factor L = 2^a 3^b into radices 8/4/2/3; returns false if L has other factors
inline bool fft_make_plan(int L, FftPlan &p) {
    p.L = L;
    p.nst = 0;
    int n = L;
    int twos = 0, threes = 0;
    while (n % 2 == 0) { n /= 2; ++twos; }
    while (n % 3 == 0) { n /= 3; ++threes; }
    if (n != 1) return false;
    while (twos >= 3) { p.rad[p.nst++] = 8; twos -= 3; }
    if (twos == 2) p.rad[p.nst++] = 4;
    if (twos == 1) p.rad[p.nst++] = 2;
    for (int i = 0; i < threes; ++i) p.rad[p.nst++] = 3;
    return p.nst <= FFT_MAX_STAGES;
}

inline bool is_2_3_smooth(int64_t n) {
    if (n < 1) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

// LDS elements to allocate for `n` used elements (swz stays inside 32-groups)
__host__ __device__ constexpr int lds_round(int n) { return (n + 31) & ~31; }

__device__ __forceinline__ int swz(int a) { return a ^ (((a >> 5) & 7) << 2); }
__device__ __forceinline__ int lds_row(int line, int e, int ls) { return swz(line * ls + e); }
__device__ __forceinline__ int lds_col(int line, int e, int nl) { return swz(e * nl + line); }

// quotient of non-negative b < 2^22 by d >= 1 through an FP32 reciprocal,
// corrected to the exact value
__device__ __forceinline__ int fdiv(int b, int d, float inv) {
    int q = (int)((float)b * inv);
    int r = b - q * d;
    if (r < 0) { --q; r += d; }
    if (r >= d) ++q;
    return q;
}

// multiply by SIGN * i
template <int SIGN, typename C> __device__ __forceinline__ C mul_si(C a) {
    C r;
    if (SIGN < 0) { r.x = a.y; r.y = -a.x; }
    else { r.x = -a.y; r.y = a.x; }
    return r;
}

template <int SIGN, typename C> __device__ __forceinline__ void dft2(C *v) {
    C a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <int SIGN, typename C> __device__ __forceinline__ void dft4(C *v) {
    C a0 = v[0] + v[2], a1 = v[0] - v[2];
    C b0 = v[1] + v[3], b1 = v[1] - v[3];
    C jb1 = mul_si<SIGN>(b1);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + jb1;
    v[3] = a1 - jb1;
}

template <int SIGN, typename C> __device__ __forceinline__ void dft3(C *v) {
    using R = decltype(C{}.x);
    const R h = (R)0.8660254037844386467637231707529362;  // sqrt(3)/2
    C t1 = v[1] + v[2];
    C t2 = v[0] - cscale(t1, (R)0.5);
    C d = cscale(v[1] - v[2], h);
    C jd = mul_si<SIGN>(d);
    v[0] = v[0] + t1;
    v[1] = t2 + jd;
    v[2] = t2 - jd;
}

template <int SIGN, typename C> __device__ __forceinline__ void dft8(C *v) {
    using R = decltype(C{}.x);
    const R s = (R)0.7071067811865475244008443621048490;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4<SIGN>(e);
    dft4<SIGN>(o);
    // w = e^{SIGN 2 pi i / 8}
    C o1, o2, o3;
    o1.x = s * (o[1].x - SIGN * o[1].y);
    o1.y = s * (o[1].y + SIGN * o[1].x);
    o2 = mul_si<SIGN>(o[2]);
    o3.x = s * (-o[3].x - SIGN * o[3].y);
    o3.y = s * (-o[3].y + SIGN * o[3].x);
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
}

template <int R, int SIGN, typename C> __device__ __forceinline__ void dftR(C *v) {
    if constexpr (R == 2) dft2<SIGN>(v);
    else if constexpr (R == 3) dft3<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v);
    else dft8<SIGN>(v);
}

template <typename T, int R, int SIGN, int NT, int MAXE, bool LFAST>
__device__ __forceinline__ void fft_stage(cpx<T> *buf, int nl, int L, int ls, int Ns,
                                          const cpx<T> *__restrict__ tw) {
    using C = cpx<T>;
    constexpr int MAXB = (MAXE + R - 1) / R;
    const int LR = L / R;
    const int nbf = nl * LR;
    const int twstep = L / (Ns * R);
    const int tid = threadIdx.x;
    const float inv_d = LFAST ? 1.0f / (float)nl : 1.0f / (float)LR;
    const float inv_ns = 1.0f / (float)Ns;
    C v[MAXB][R];
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        const int b = tid + t * NT;
        if (b < nbf) {
            int line, j;
            if (LFAST) { j = fdiv(b, nl, inv_d); line = b - j * nl; }
            else { line = fdiv(b, LR, inv_d); j = b - line * LR; }
#pragma unroll
            for (int r = 0; r < R; ++r)
                v[t][r] = buf[LFAST ? lds_col(line, j + r * LR, nl) : lds_row(line, j + r * LR, ls)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        const int b = tid + t * NT;
        if (b < nbf) {
            int line, j;
            if (LFAST) { j = fdiv(b, nl, inv_d); line = b - j * nl; }
            else { line = fdiv(b, LR, inv_d); j = b - line * LR; }
            const int k = j - fdiv(j, Ns, inv_ns) * Ns;
            if (k != 0) {
                // powers of w^k from one (radix 8: two) table loads
                C w[R];
                w[1] = tw[k * twstep];
                if (SIGN > 0) w[1].y = -w[1].y;
                if constexpr (R >= 3) w[2] = cmul(w[1], w[1]);
                if constexpr (R >= 4) w[3] = cmul(w[2], w[1]);
                if constexpr (R == 8) {
                    w[4] = tw[4 * k * twstep];
                    if (SIGN > 0) w[4].y = -w[4].y;
                    w[5] = cmul(w[4], w[1]);
                    w[6] = cmul(w[4], w[2]);
                    w[7] = cmul(w[4], w[3]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], w[r]);
            }
            dftR<R, SIGN>(v[t]);
            const int o = (j - k) * R + k;   // (j/Ns)*Ns*R + k
#pragma unroll
            for (int r = 0; r < R; ++r)
                buf[LFAST ? lds_col(line, o + r * Ns, nl) : lds_row(line, o + r * Ns, ls)] = v[t][r];
        }
    }
    __syncthreads();
}

// Full transform of nl lines (all threads of the block must call this).
// Requires nl * L <= NT * MAXE.
template <typename T, int SIGN, int NT, int MAXE, bool LFAST>
__device__ void fft_lds(cpx<T> *buf, int nl, int ls, const FftPlan &plan,
                        const cpx<T> *__restrict__ tw) {
    int Ns = 1;
    const int L = plan.L;
    for (int s = 0; s < plan.nst; ++s) {
        const int R = plan.rad[s];
        switch (R) {
        case 8: fft_stage<T, 8, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        case 4: fft_stage<T, 4, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        case 3: fft_stage<T, 3, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        default: fft_stage<T, 2, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        }
        Ns *= R;
    }
}

// FP64 twiddle generation: tw[k] = exp(-2 pi i k / L)
template <typename T>
__global__ void k_twiddles(cpx<T> *tw, int L) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L) {
        double s, c;
        sincospi(-2.0 * (double)k / (double)L, &s, &c);
        cpx<T> z;
        z.x = (T)c;
        z.y = (T)s;
        tw[k] = z;
    }
}

} // namespace lc

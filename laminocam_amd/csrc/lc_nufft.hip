// 2-D polar-grid NUFFT projectors on gfx950 (forward2d / backproj2d,
// src/projection.cpp:146-182).
//
// Type 1 (backproj2d):  f_m = sum_j c_j exp(+i m.w_j),   w_j = r_k (cos th_i, sin th_i)
// Type 2 (forward2d):   c_j = sum_m f_m exp(-i m.w_j)
// with modes m in [-N/2, N/2) (CMCL order; include/nufft.h:64-95).  The
// method is the exponential-of-semicircle NUFFT (Barnett, Magland and
// af Klinteberg, SIAM J. Sci. Comput. 41(5), 2019): spread / interpolate
// with psi(u) = exp(beta (sqrt(1 - (2u/w)^2) - 1)) on a fine grid of nf = 2N
// (rounded to 2^p 3^q) cells, FFT the fine grid, divide by the kernel's
// Fourier transform.  The fine grid is indexed with its centre at nf/2 so a
// radial line of points is a straight segment of cells; the resulting
// (-1)^m phase is folded into the correction factors.
//
// Spreading is deterministic: each workgroup owns a T x T tile of the fine
// grid, gathers the points whose footprint reaches the tile (ray / box
// intersection per angle, angles tested 64 at a time), and each of its waves
// accumulates its own share of the rays into a private LDS copy of the tile,
// one point at a time with the 8 x 8 footprint spread over the 64 lanes.  The
// copies are summed in a fixed order at the end.
#include <cmath>

#include <rocprim/device/device_radix_sort.hpp>

#include "lc_engine.hpp"


namespace lc {

extern __shared__ __align__(16) char lc_smem[];

constexpr int NU_W = 8;        // kernel width (cells); 64 lanes cover the footprint
// spread tile edge (cells): 16 x 16 tiles (2 KiB per wave copy) keep more
// tiles in flight than 32 x 32; faster for single slices and on the served
// stream, where back-projections stack a few slices per geometry
constexpr int NU_T = 16;
constexpr int NU_WAVES = 4;    // waves per spread workgroup
static_assert(NU_T == 16 || NU_T == 32, "spread tile rows of 16 (swizzled) or 32 cells");

// spread tile cell (y, x) -> LDS index.  For rows of 16 cells x is XOR-ed with
// bit 1 of y (times 8), so the 4 rows x 8 columns a half-wave updates fall on
// distinct bank pairs (plain rows y and y + 2 would share banks).
__device__ __forceinline__ int nu_swx(int y) { return NU_T == 16 ? (((y >> 1) & 1) << 3) : 0; }
__device__ __forceinline__ int nu_tidx(int y, int x) { return y * NU_T + (x ^ nu_swx(y)); }

// value of v in lane `lane` (uniform lane index: a scalar broadcast, no LDS)
__device__ __forceinline__ float rl_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double rl_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// exp(beta (sqrt(1 - z^2) - 1)) for |z| < 1, else 0.  The hardware square
// root (within 1 ulp; s lies in (0, 1]) instead of the correctly rounded
// sequence: one instruction instead of about fifteen, a relative weight
// error near 1e-6, inside the kernel's own approximation error
__device__ __forceinline__ float es_kernel(float z, float beta) {
    const float s = 1.0f - z * z;
    return s > 0.0f ? __expf(beta * (__builtin_amdgcn_sqrtf(s) - 1.0f)) : 0.0f;
}

// ---------------------------------------------------------------------------
// detector transforms along each sinogram row (length n = nrad)
// backproj: G = ifftshift1(fft1(fftshift1(y)))      (complex out)
// forward:  y = Re(fftshift1(ifft1(ifftshift1(C)))) / n
// fftshift: out[k] = in[(k + n/2) % n]; ifftshift: d = n/2 + (n odd)
// ---------------------------------------------------------------------------
// LC > 0 (every kernel of this file): line length LC compiled, with unrolled
// stages (fft_lds_ct); 0: the runtime plan
template <int SIGN, int NT, int MAXE, bool LFAST, int LC>
__device__ __forceinline__ void nu_fft(float2 *buf, int nl, int L, int lsh, const FftPlan &plan,
                                       const float2 *__restrict__ tw) {
    if constexpr (LC > 0) fft_lds_ct<float, LC, SIGN, NT, MAXE, LFAST>(buf, nl, lsh, tw);
    else fft_lds<float, SIGN, NT, MAXE, LFAST>(buf, nl, LFAST ? 0 : L, plan, tw);
}

template <int SIGN, int NT, int MAXE, int LC>
__global__ void __launch_bounds__(NT)
k_det_fft(const float *__restrict__ real_in, const float2 *__restrict__ cplx_in,
          float2 *__restrict__ cplx_out, float *__restrict__ real_out, int nrows, int n, int nl,
          FftPlan plan, const float2 *__restrict__ tw, float scale) {
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    if constexpr (LC > 0) n = LC;
    // blocks never straddle slices (blockIdx.y = slice, nrows rows each): a
    // row's place in its block, which its rounding depends on, is then the
    // same whether the slice is transformed alone or inside a stack
    const int lrow0 = blockIdx.x * nl;
    const int nloc = min(nl, nrows - lrow0);
    const int64_t row0 = (int64_t)blockIdx.y * nrows + lrow0;
    const int dfw = n / 2, dinv = n / 2 + (n & 1);
    // input shift: fftshift for backproj (SIGN -1), ifftshift for forward (+1)
    const int din = SIGN < 0 ? dfw : dinv;
    const int dout = SIGN < 0 ? dinv : dfw;
    for (int idx = threadIdx.x; idx < nl * n; idx += NT) {
        const int l = idx / n, k = idx - l * n;
        float2 v = make_float2(0.f, 0.f);
        if (l < nloc) {
            int src = k + din;
            if (src >= n) src -= n;
            const int64_t q = (row0 + l) * n + src;
            v = SIGN < 0 ? make_float2(real_in[q], 0.f) : cplx_in[q];
        }
        buf[lds_row(l, k, n)] = v;
    }
    __syncthreads();
    nu_fft<SIGN, NT, MAXE, false, LC>(buf, nl, n, 0, plan, tw);
    for (int idx = threadIdx.x; idx < nloc * n; idx += NT) {
        const int l = idx / n, k = idx - l * n;
        int src = k + dout;
        if (src >= n) src -= n;
        const float2 v = buf[lds_row(l, src, n)];
        const int64_t q = (row0 + l) * n + k;
        if (SIGN < 0) cplx_out[q] = v;
        else real_out[q] = v.x * scale;
    }
}

// direct DFT fallback for detector widths that are not 2^p 3^q
template <int SIGN>
__global__ void k_det_dft(const float *__restrict__ real_in, const float2 *__restrict__ cplx_in,
                          float2 *__restrict__ cplx_out, float *__restrict__ real_out, int nrows,
                          int n, const float2 *__restrict__ tw, float scale) {
    const int row = blockIdx.x;   // grid: one block per row
    const int dfw = n / 2, dinv = n / 2 + (n & 1);
    const int din = SIGN < 0 ? dfw : dinv, dout = SIGN < 0 ? dinv : dfw;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        // output index k reads transform bin kk
        int kk = k + dout;
        if (kk >= n) kk -= n;
        double re = 0.0, im = 0.0;
        for (int j = 0; j < n; ++j) {
            int src = j + din;
            if (src >= n) src -= n;
            const int64_t q = (int64_t)row * n + src;
            const float2 v = SIGN < 0 ? make_float2(real_in[q], 0.f) : cplx_in[q];
            float2 w = tw[(int)(((int64_t)j * kk) % n)];
            if (SIGN > 0) w.y = -w.y;
            re += (double)v.x * w.x - (double)v.y * w.y;
            im += (double)v.x * w.y + (double)v.y * w.x;
        }
        const int64_t q = (int64_t)row * n + k;
        if (SIGN < 0) cplx_out[q] = make_float2((float)re, (float)im);
        else real_out[q] = (float)re * scale;
    }
}

// ---------------------------------------------------------------------------
// Type-1 spreading (deterministic tile gather)
// ---------------------------------------------------------------------------
struct SpreadGeom {
    int nang, nrad, nf, ntx;   // ntx tiles per fine-grid edge
    double dr;                 // 2 pi / nrad
    double sc;                 // nf / (2 pi): radians -> cells
    float beta;
};

// intersect the ray of angle (c, s) with box [x0,x1]x[y0,y1] (cells, shifted
// grid) -> radial index range [k0, k1] (empty if k0 > k1)
__device__ __forceinline__ void ray_box(double c, double s, const SpreadGeom &g, double x0,
                                        double x1, double y0, double y1, int &k0, int &k1) {
    // g(k) = nf/2 + sc * r_k * (c, s),  r_k = (k + 1/2) dr - pi
    double rlo = -1e300, rhi = 1e300;
    const double h = 0.5 * g.nf;
    const double ax = g.sc * c, ay = g.sc * s;
    if (fabs(ax) > 1e-12) {
        double a = (x0 - h) / ax, b = (x1 - h) / ax;
        if (a > b) { double t = a; a = b; b = t; }
        rlo = fmax(rlo, a);
        rhi = fmin(rhi, b);
    } else if (h < x0 || h > x1) {
        k0 = 1;
        k1 = 0;
        return;
    }
    if (fabs(ay) > 1e-12) {
        double a = (y0 - h) / ay, b = (y1 - h) / ay;
        if (a > b) { double t = a; a = b; b = t; }
        rlo = fmax(rlo, a);
        rhi = fmin(rhi, b);
    } else if (h < y0 || h > y1) {
        k0 = 1;
        k1 = 0;
        return;
    }
    if (rlo > rhi) {
        k0 = 1;
        k1 = 0;
        return;
    }
    // r = (k + 1/2) dr - pi  ->  k = (r + pi)/dr - 1/2 ; widen by one
    const double klo = (rlo + M_PI) / g.dr - 0.5, khi = (rhi + M_PI) / g.dr - 0.5;
    k0 = max(0, (int)floor(klo) - 1);
    k1 = min(g.nrad - 1, (int)ceil(khi) + 1);
}

// grid: (ntx * ntx, nslices).  C: per slice nang x nrad complex (detector-
// transformed sinogram).  fine: per slice nf x nf complex.
__global__ void __launch_bounds__(64 * NU_WAVES)
k_spread2d(const float2 *__restrict__ C, float2 *__restrict__ fine, const double2 *__restrict__ cs,
           SpreadGeom g) {
    float2 *acc = reinterpret_cast<float2 *>(lc_smem);   // [NU_WAVES][T*T]
    // [NU_WAVES][64][2 NU_W]: kernel weights of the wave's current 64 points
    float *wtab = reinterpret_cast<float *>(acc + NU_WAVES * NU_T * NU_T);
    const int slice = blockIdx.y;
    const int tile = blockIdx.x;
    const int ty = tile / g.ntx, tx = tile - ty * g.ntx;
    const int x0 = tx * NU_T, y0 = ty * NU_T;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float2 *my = acc + wave * NU_T * NU_T;
    float *wt = wtab + wave * 64 * 2 * NU_W;
    for (int i = lane; i < NU_T * NU_T; i += 64) my[i] = make_float2(0.f, 0.f);
    __syncthreads();
    const float2 *Cs = C + (int64_t)slice * g.nang * g.nrad;
    const int dx = lane & 7, dy = lane >> 3;
    const double hw = 0.5 * NU_W;
    const float inv_hw = 1.0f / (float)hw;
    // periodic images of the tile box that can receive wrapped contributions
    for (int oy = -1; oy <= 1; ++oy) {
        for (int ox = -1; ox <= 1; ++ox) {
            const double bx0 = x0 + ox * g.nf - hw - 1.0, bx1 = x0 + NU_T - 1 + ox * g.nf + hw + 1.0;
            const double by0 = y0 + oy * g.nf - hw - 1.0, by1 = y0 + NU_T - 1 + oy * g.nf + hw + 1.0;
            if (bx1 < 0.0 || bx0 > g.nf || by1 < 0.0 || by0 > g.nf) continue;
            // waves take 64-angle chunks round robin; lanes test one angle each
            for (int a0 = wave * 64; a0 < g.nang; a0 += 64 * NU_WAVES) {
                const int ai = a0 + lane;
                int k0 = 1, k1 = 0;
                double cth = 0.0, sth = 0.0;
                if (ai < g.nang) {
                    const double2 t = cs[ai];
                    cth = t.x;
                    sth = t.y;
                    ray_box(cth, sth, g, bx0, bx1, by0, by1, k0, k1);
                }
                // samples of the hit rays packed 64 to a batch, in ray order
                // then ascending k (the order a ray-by-ray walk spreads
                // them): lane q of the packed sequence finds its ray by a
                // binary search over the wave's inclusive prefix sums
                const int nk = k0 <= k1 ? k1 - k0 + 1 : 0;
                int incl = nk;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += t;
                }
                const int total = __shfl(incl, 63, 64);
                for (int base = 0; base < total; base += 64) {
                    {
                        const int q = base + lane;
                        int rr = 0;
#pragma unroll
                        for (int step = 32; step > 0; step >>= 1)
                            if (__shfl(incl, rr + step - 1, 64) <= q) rr += step;
                        rr = rr < 64 ? rr : 63;
                        const int rincl = __shfl(incl, rr, 64), rnk = __shfl(nk, rr, 64);
                        const int rk0 = __shfl(k0, rr, 64);
                        const double rc = __shfl(cth, rr, 64), rs = __shfl(sth, rr, 64);
                        const int k = rk0 + (q - (rincl - rnk));
                        const int ang = a0 + rr;
                        int lx0 = 0, ly0 = 0;
                        float fx = 0.f, fy = 0.f;
                        float2 cv = make_float2(0.f, 0.f);
                        bool ok = q < total;
                        if (ok) {
                            const double r = ((double)k + 0.5) * g.dr - M_PI;
                            const double gx = 0.5 * g.nf + g.sc * r * rc - ox * g.nf;
                            const double gy = 0.5 * g.nf + g.sc * r * rs - oy * g.nf;
                            lx0 = (int)ceil(gx - hw);
                            ly0 = (int)ceil(gy - hw);
                            fx = (float)((double)lx0 - gx);   // footprint start - point, in (-hw - 1, -hw]
                            fy = (float)((double)ly0 - gy);
                            ok = lx0 + NU_W - 1 >= x0 && lx0 < x0 + NU_T && ly0 + NU_W - 1 >= y0 &&
                                 ly0 < y0 + NU_T;
                            if (ok) cv = Cs[(int64_t)ang * g.nrad + k];
                        }
                        // the point's 2 x 8 kernel weights, once per point (not
                        // per footprint cell): the lanes of the spread loop read
                        // them back from the wave's table
                        wave_lds_fence();
                        if (ok) {
#pragma unroll
                            for (int d = 0; d < NU_W; ++d) {
                                wt[lane * 2 * NU_W + d] = es_kernel(((float)d + fx) * inv_hw, g.beta);
                                wt[lane * 2 * NU_W + NU_W + d] = es_kernel(((float)d + fy) * inv_hw, g.beta);
                            }
                        }
                        wave_lds_fence();
                        unsigned long long pts = __ballot(ok);
                        while (pts) {
                            const int s2 = __ffsll((long long)pts) - 1;
                            pts &= pts - 1;
                            const int px0 = __builtin_amdgcn_readlane(lx0, s2);
                            const int py0 = __builtin_amdgcn_readlane(ly0, s2);
                            const int cx = px0 + dx - x0, cy = py0 + dy - y0;
                            if (cx >= 0 && cx < NU_T && cy >= 0 && cy < NU_T) {
                                const float wgt = wt[s2 * 2 * NU_W + dx] * wt[s2 * 2 * NU_W + NU_W + dy];
                                const float vr = rl_f32(cv.x, s2), vi = rl_f32(cv.y, s2);
                                float2 &acc1 = my[nu_tidx(cy, cx)];
                                acc1.x += wgt * vr;
                                acc1.y += wgt * vi;
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // fixed-order sum of the wave copies; each tile cell has one writer
    float2 *fs = fine + (int64_t)slice * g.nf * g.nf;
    for (int i = threadIdx.x; i < NU_T * NU_T; i += 64 * NU_WAVES) {
        float2 s = acc[i];
        for (int w = 1; w < NU_WAVES; ++w) {
            s.x += acc[w * NU_T * NU_T + i].x;
            s.y += acc[w * NU_T * NU_T + i].y;
        }
        const int cy = i / NU_T, cx = (i - cy * NU_T) ^ nu_swx(cy);
        const int gxx = x0 + cx, gyy = y0 + cy;
        if (gxx < g.nf && gyy < g.nf) fs[(int64_t)gyy * g.nf + gxx] = s;
    }
}

// ---------------------------------------------------------------------------
// Type-1 spreading by sorted footprint-corner bins (the form of the 3-D
// spread).  The fine grid of every slice is cut into B2 x B2 bins of
// footprint corners (a point's corner is its first footprint cell,
// ceil(g - 4) per axis, taken modulo nf).  k_bin2_keys gives every point of
// the stack its (slice, bin) key; a stable radix sort orders the points by
// key, keeping point order inside a bin; k_bin2_ranges marks each bin's run.
// A wave then owns one bin: it takes the bin's points in order and spreads
// each point's 8 x 8 footprint (one cell per lane) into its LDS copy of the
// bin's padded region (B2 + 7 cells per edge), then adds the region into the
// slice's fine grid, wrapping the overhang.  When nf is a multiple of 2 B2
// the bins of one parity per axis have disjoint wrapped regions, so the bins
// run in 4 phases, each adding without conflicts, fullest bins first.  Every
// sum runs in a fixed order and a slice's points and bins do not depend on
// the other slices of the stack: the result is deterministic and a slice
// spreads the same alone as in a stack.
// ---------------------------------------------------------------------------
// A bin shape the spread is compiled for: bin edge B corner cells, padded
// region edge R = B + 7, region row stride RS = B + 8 (rows y and y + 1 fall
// on opposite bank halves), heavy-bin threshold (points) and waves (region
// copies within 64 KiB).  The autotuner picks one per fine-grid size
// (spread2d_autotune); shape 0 (B = 16) is the default.
template <int B_> struct Shape2 {
    static constexpr int B = B_, R = B + NU_W - 1, RS = B + NU_W;
    static constexpr int RC = R * RS;                 // cells of one region copy
    static constexpr unsigned HEAVY = 2u * B * B;      // 512 points for B = 16
    static constexpr int NWH = 65536 / (8 * RC) > 12 ? 12 : 65536 / (8 * RC);
    static_assert(RC % 2 == 0, "the region clears as float4");
};
constexpr int N_SHAPES2 = 2;
constexpr int shape2_edge(int i) { return i == 1 ? 32 : 16; }
template <typename F> void with_shape2(int i, F &&f) {
    if (i == 1) f(Shape2<32>{});
    else f(Shape2<16>{});
}
constexpr int B2_ORDER_BITS = 28;      // launch-order key: phase (2 bits) above a 26-bit count

struct Bins2 {
    int nb;             // bins per fine-grid edge (nf / B)
    unsigned nbins;     // bins per slice
    int npts;           // points per slice (nang * nrad)
};

// fine-grid footprint corner (unwrapped) and its offsets of point p of a slice
__device__ __forceinline__ void corner2(const double2 *__restrict__ cs, const SpreadGeom &g, int p, int &lx0,
                                        int &ly0, float &fx, float &fy) {
    const int ang = p / g.nrad, k = p - ang * g.nrad;
    const double2 t = cs[ang];
    const double r = ((double)k + 0.5) * g.dr - M_PI;
    const double gx = 0.5 * g.nf + g.sc * r * t.x, gy = 0.5 * g.nf + g.sc * r * t.y;
    const double hw = 0.5 * NU_W;
    lx0 = (int)ceil(gx - hw);
    ly0 = (int)ceil(gy - hw);
    fx = (float)((double)lx0 - gx);   // footprint start - point, in (-hw - 1, -hw]
    fy = (float)((double)ly0 - gy);
}

__device__ __forceinline__ int wrap_nf(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}

// key[i] = slice * nbins + bin of point i of the stack, val[i] = i
template <typename S>
__global__ void k_bin2_keys(const double2 *__restrict__ cs, SpreadGeom g, Bins2 bn, unsigned *__restrict__ key,
                            unsigned *__restrict__ val, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = (int)(i / bn.npts), p = (int)(i - (int64_t)s * bn.npts);
    int lx0, ly0;
    float fx, fy;
    corner2(cs, g, p, lx0, ly0, fx, fy);
    const int bx = wrap_nf(lx0, g.nf) / S::B, by = wrap_nf(ly0, g.nf) / S::B;
    key[i] = (unsigned)s * bn.nbins + (unsigned)(by * bn.nb + bx);
    val[i] = (unsigned)i;
}

// [start[b], end[b]) is bin b's run in the sorted keys (both cleared first)
__global__ void k_bin2_ranges(const unsigned *__restrict__ skey, int64_t n, unsigned *__restrict__ start,
                              unsigned *__restrict__ end) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned b = skey[i];
    if (i == 0 || skey[i - 1] != b) start[b] = (unsigned)i;
    if (i == n - 1 || skey[i + 1] != b) end[b] = (unsigned)(i + 1);
}

// launch order: phase (bin parity per axis), then fullest first
__global__ void k_bin2_order_keys(const unsigned *__restrict__ bstart, const unsigned *__restrict__ bend, Bins2 bn,
                                  unsigned *__restrict__ key, unsigned *__restrict__ val, unsigned total) {
    const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= total) return;
    const unsigned local = b % bn.nbins;
    const unsigned by = local / bn.nb, bx = local - by * bn.nb;
    const unsigned ph = ((by & 1) << 1) | (bx & 1);
    const unsigned cnt = min(bend[b] - bstart[b], (1u << 26) - 1);
    key[b] = (ph << 26) | ((1u << 26) - 1 - cnt);
    val[b] = b;
}

__global__ void k_zero_u32_2(unsigned *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_zero_f2(float2 *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_float2(0.f, 0.f);
}

// Bins of more than S::HEAVY points (near the k-space origin every angle
// passes: ~8 samples per angle in the central 16 x 16 bin) are spread by
// S::NWH waves, each taking a consecutive part of the bin's points into its
// own copy of the region; the copies are summed in wave order.  Whether a bin
// is heavy depends only on its own count, so a slice's result is the same
// alone and in a stack.

// spread bin bid's points with the block's NW waves into their region copies
// (reg: NW x R2 x R2S) and add the copies' fixed-order sum into the fine grid
template <typename S, int NW>
__device__ __forceinline__ void spread2_bin(const float2 *__restrict__ C, float2 *__restrict__ fine,
                                            const double2 *__restrict__ cs, const unsigned *__restrict__ sval,
                                            unsigned bid, unsigned s0, unsigned s1, const SpreadGeom &g,
                                            const Bins2 &bn, float2 *reg) {
    const int slice = (int)(bid / bn.nbins);
    const int local = (int)(bid - (unsigned)slice * bn.nbins);
    const int by = local / bn.nb, bx = local - by * bn.nb;
    constexpr int B2 = S::B, R2 = S::R, R2S = S::RS;
    const int c0y = by * B2, c0x = bx * B2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int RC = S::RC;
    for (int i = threadIdx.x; i < NW * RC / 2; i += 64 * NW)
        reinterpret_cast<float4 *>(reg)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    float2 *my = reg + wave * RC;
    const int dx = lane & 7, dy = lane >> 3;
    const float inv_hw = 1.0f / (0.5f * NU_W);
    const float2 *Cs = C + (int64_t)slice * bn.npts;
    const unsigned pbase = (unsigned)slice * (unsigned)bn.npts;
    // wave w takes the w-th part of the bin's points, in order
    const unsigned per = (s1 - s0 + NW - 1) / NW;
    const unsigned ws = s0 + wave * per, we = min(s1, ws + per);
    for (unsigned base = ws; base < we; base += 64) {
        const unsigned i = base + lane;
        const bool ok = i < we;
        int tx = 0, ty = 0;
        float fx = 0.f, fy = 0.f;
        float2 cv = make_float2(0.f, 0.f);
        if (ok) {
            const int p = (int)(sval[i] - pbase);
            int lx0, ly0;
            corner2(cs, g, p, lx0, ly0, fx, fy);
            // the corner taken modulo nf, relative to the bin: in [0, B2)
            tx = wrap_nf(lx0, g.nf) - c0x;
            ty = wrap_nf(ly0, g.nf) - c0y;
            cv = Cs[p];
        }
        unsigned long long pts = __ballot(ok);
        while (pts) {
            const int s2 = __ffsll((long long)pts) - 1;
            pts &= pts - 1;
            const int px0 = __builtin_amdgcn_readlane(tx, s2);
            const int py0 = __builtin_amdgcn_readlane(ty, s2);
            const float wgt = es_kernel(((float)dx + rl_f32(fx, s2)) * inv_hw, g.beta) *
                              es_kernel(((float)dy + rl_f32(fy, s2)) * inv_hw, g.beta);
            const float vr = rl_f32(cv.x, s2), vi = rl_f32(cv.y, s2);
            float2 &a = my[(py0 + dy) * R2S + px0 + dx];
            a.x += wgt * vr;
            a.y += wgt * vi;
        }
    }
    __syncthreads();
    float2 *fs = fine + (int64_t)slice * g.nf * g.nf;
    for (int i = threadIdx.x; i < R2 * R2; i += 64 * NW) {
        const int y = i / R2, x = i - y * R2;
        int gy = c0y + y, gx = c0x + x;
        gy = gy < g.nf ? gy : gy - g.nf;
        gx = gx < g.nf ? gx : gx - g.nf;
        float2 v = reg[y * R2S + x];
        for (int w = 1; w < NW; ++w) {
            v.x += reg[w * RC + y * R2S + x].x;
            v.y += reg[w * RC + y * R2S + x].y;
        }
        float2 &f = fs[(int64_t)gy * g.nf + gx];
        f.x += v.x;
        f.y += v.y;
    }
}

// one wave per light bin of this phase (order: the phase's bins, fullest first)
template <typename S>
__global__ void __launch_bounds__(64)
k_spread2d_bin(const float2 *__restrict__ C, float2 *__restrict__ fine, const double2 *__restrict__ cs,
               const unsigned *__restrict__ sval, const unsigned *__restrict__ bstart,
               const unsigned *__restrict__ bend, const unsigned *__restrict__ order, SpreadGeom g, Bins2 bn) {
    const unsigned bid = order[blockIdx.x];
    const unsigned s0 = bstart[bid], s1 = bend[bid];
    if (s0 >= s1 || s1 - s0 > S::HEAVY) return;   // empty, or a heavy bin
    spread2_bin<S, 1>(C, fine, cs, sval, bid, s0, s1, g, bn, reinterpret_cast<float2 *>(lc_smem));
}

// the phase's heavy bins (a prefix of its order): each block takes bins
// j, j + gridDim.x, ... until the first light one
template <typename S>
__global__ void __launch_bounds__(64 * S::NWH)
k_spread2d_bin_heavy(const float2 *__restrict__ C, float2 *__restrict__ fine, const double2 *__restrict__ cs,
                     const unsigned *__restrict__ sval, const unsigned *__restrict__ bstart,
                     const unsigned *__restrict__ bend, const unsigned *__restrict__ order, unsigned nphase,
                     SpreadGeom g, Bins2 bn) {
    for (unsigned j = blockIdx.x; j < nphase; j += gridDim.x) {
        const unsigned bid = order[j];
        const unsigned s0 = bstart[bid], s1 = bend[bid];
        if (s1 - s0 <= S::HEAVY) break;   // the order is fullest first: the rest are light
        spread2_bin<S, S::NWH>(C, fine, cs, sval, bid, s0, s1, g, bn, reinterpret_cast<float2 *>(lc_smem));
        __syncthreads();   // the copies are cleared for the next bin only after every wave has flushed
    }
}

// ---------------------------------------------------------------------------
// Type-2 interpolation: one thread per point
// ---------------------------------------------------------------------------
__global__ void k_interp2d(const float2 *__restrict__ fine, float2 *__restrict__ C,
                           const double2 *__restrict__ cs, SpreadGeom g, int npts) {
    const int slice = blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int ang = p / g.nrad, k = p - ang * g.nrad;
    const double c = cs[ang].x, s = cs[ang].y;
    const double r = ((double)k + 0.5) * g.dr - M_PI;
    const double gx = 0.5 * g.nf + g.sc * r * c, gy = 0.5 * g.nf + g.sc * r * s;
    const double hw = 0.5 * NU_W;
    const int lx0 = (int)ceil(gx - hw), ly0 = (int)ceil(gy - hw);
    const float inv_hw = 1.0f / (float)hw;
    float wx[NU_W], wy[NU_W];
#pragma unroll
    for (int d = 0; d < NU_W; ++d) {
        wx[d] = es_kernel((float)((double)(lx0 + d) - gx) * inv_hw, g.beta);
        wy[d] = es_kernel((float)((double)(ly0 + d) - gy) * inv_hw, g.beta);
    }
    const float2 *fs = fine + (int64_t)slice * g.nf * g.nf;
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int a = 0; a < NU_W; ++a) {
        int yy = ly0 + a;
        yy = yy < 0 ? yy + g.nf : (yy >= g.nf ? yy - g.nf : yy);
        const float2 *row = fs + (int64_t)yy * g.nf;
        float rr = 0.f, ri = 0.f;
#pragma unroll
        for (int b = 0; b < NU_W; ++b) {
            int xx = lx0 + b;
            xx = xx < 0 ? xx + g.nf : (xx >= g.nf ? xx - g.nf : xx);
            const float2 v = row[xx];
            rr += wx[b] * v.x;
            ri += wx[b] * v.y;
        }
        re += wy[a] * rr;
        im += wy[a] * ri;
    }
    C[(int64_t)slice * npts + p] = make_float2(re, im);
}

// ---------------------------------------------------------------------------
// fine-grid transforms (pruned to the N x N modes)
// fine row r, column l <-> cell (x = l, y = r); mode m <-> fine index m mod nf
// V: nf rows x N mode columns (row major), per slice
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mode_to_fine(int a, int N, int nf) {
    const int m = a - N / 2;
    return m < 0 ? m + nf : m;
}

// type 1, pass A: FFT+ of every fine row, keep the N mode columns -> V
template <int NT, int MAXE, int LC>
__global__ void __launch_bounds__(NT)
k_fine_rows_t1(const float2 *__restrict__ fine, float2 *__restrict__ V, int nf, int N, int nl,
               FftPlan plan, const float2 *__restrict__ tw) {
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    if constexpr (LC > 0) nf = LC;
    const int slice = blockIdx.y;
    const int row0 = blockIdx.x * nl;
    const int nloc = min(nl, nf - row0);
    const float2 *fs = fine + (int64_t)slice * nf * nf;
    for (int idx = threadIdx.x; idx < nl * nf; idx += NT) {
        const int l = idx / nf, e = idx - l * nf;
        buf[lds_row(l, e, nf)] = l < nloc ? fs[(int64_t)(row0 + l) * nf + e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    nu_fft<1, NT, MAXE, false, LC>(buf, nl, nf, 0, plan, tw);
    float2 *vs = V + (int64_t)slice * nf * N;
    for (int idx = threadIdx.x; idx < nloc * N; idx += NT) {
        const int l = idx / N, b = idx - l * N;
        vs[(int64_t)(row0 + l) * N + b] = buf[lds_row(l, mode_to_fine(b, N, nf), nf)];
    }
}

// type 1, pass B: FFT+ down each mode column, keep the N mode rows,
// deconvolve, real part / nrad -> image (N x N real)
// LC > 0: B = 1 << lsh
template <int NT, int MAXE, int LC>
__global__ void __launch_bounds__(NT)
k_fine_cols_t1(const float2 *__restrict__ V, float *__restrict__ img, const float *__restrict__ corr,
               int nf, int N, int B, int lsh, float scale, FftPlan plan, const float2 *__restrict__ tw) {
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    if constexpr (LC > 0) nf = LC;
    const int slice = blockIdx.y;
    const int col0 = blockIdx.x * B;
    const float2 *vs = V + (int64_t)slice * nf * N;
    for (int idx = threadIdx.x; idx < nf * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        buf[lds_col(c, i, B)] = col0 + c < N ? vs[(int64_t)i * N + col0 + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    nu_fft<1, NT, MAXE, true, LC>(buf, B, nf, lsh, plan, tw);
    float *is = img + (int64_t)slice * N * N;
    for (int idx = threadIdx.x; idx < N * B; idx += NT) {
        const int a = idx / B, c = idx - a * B;
        const int b = col0 + c;
        if (b < N) is[(int64_t)a * N + b] = buf[lds_col(c, mode_to_fine(a, N, nf), B)].x * corr[a] * corr[b] * scale;
    }
}

// type 2, pass B: deconvolve the image, place the N mode rows on the fine
// column, FFT- -> V (nf x N)
template <int NT, int MAXE, int LC>
__global__ void __launch_bounds__(NT)
k_fine_cols_t2(const float *__restrict__ img, float2 *__restrict__ V, const float *__restrict__ corr,
               int nf, int N, int B, int lsh, FftPlan plan, const float2 *__restrict__ tw) {
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    if constexpr (LC > 0) nf = LC;
    const int slice = blockIdx.y;
    const int col0 = blockIdx.x * B;
    const float *is = img + (int64_t)slice * N * N;
    for (int idx = threadIdx.x; idx < nf * B; idx += NT) buf[swz(idx)] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int idx = threadIdx.x; idx < N * B; idx += NT) {
        const int a = idx / B, c = idx - a * B;
        const int b = col0 + c;
        if (b < N) buf[lds_col(c, mode_to_fine(a, N, nf), B)] = make_float2(is[(int64_t)a * N + b] * corr[a] * corr[b], 0.f);
    }
    __syncthreads();
    nu_fft<-1, NT, MAXE, true, LC>(buf, B, nf, lsh, plan, tw);
    float2 *vs = V + (int64_t)slice * nf * N;
    for (int idx = threadIdx.x; idx < nf * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        if (col0 + c < N) vs[(int64_t)i * N + col0 + c] = buf[lds_col(c, i, B)];
    }
}

// type 2, pass A: place the N mode columns of each V row on the fine row,
// FFT- -> fine grid
template <int NT, int MAXE, int LC>
__global__ void __launch_bounds__(NT)
k_fine_rows_t2(const float2 *__restrict__ V, float2 *__restrict__ fine, int nf, int N, int nl,
               FftPlan plan, const float2 *__restrict__ tw) {
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    if constexpr (LC > 0) nf = LC;
    const int slice = blockIdx.y;
    const int row0 = blockIdx.x * nl;
    const int nloc = min(nl, nf - row0);
    for (int idx = threadIdx.x; idx < nl * nf; idx += NT) buf[swz(idx)] = make_float2(0.f, 0.f);
    __syncthreads();
    const float2 *vs = V + (int64_t)slice * nf * N;
    for (int idx = threadIdx.x; idx < nloc * N; idx += NT) {
        const int l = idx / N, b = idx - l * N;
        buf[lds_row(l, mode_to_fine(b, N, nf), nf)] = vs[(int64_t)(row0 + l) * N + b];
    }
    __syncthreads();
    nu_fft<-1, NT, MAXE, false, LC>(buf, nl, nf, 0, plan, tw);
    float2 *fs = fine + (int64_t)slice * nf * nf;
    for (int idx = threadIdx.x; idx < nloc * nf; idx += NT) {
        const int l = idx / nf, e = idx - l * nf;
        fs[(int64_t)(row0 + l) * nf + e] = buf[lds_row(l, e, nf)];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Gauss-Legendre quadrature of the kernel's Fourier transform (FP64):
// Psi(xi) = (w/2) int_{-pi/2}^{pi/2} e^{beta (cos t - 1)} cos(xi (w/2) sin t) cos t dt
static double es_phihat(double beta, int w, double xi) {
    static double gx[96], gw[96];
    static bool ready = false;
    const int n = 96;
    if (!ready) {
        for (int i = 0; i < n; ++i) {
            double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = x;
                for (int k = 2; k <= n; ++k) {
                    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                double dp = n * (x * p1 - p0) / (x * x - 1.0);
                double dx = p1 / dp;
                x -= dx;
                if (std::fabs(dx) < 1e-16) break;
            }
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double dp = n * (x * p1 - p0) / (x * x - 1.0);
            gx[i] = x;
            gw[i] = 2.0 / ((1.0 - x * x) * dp * dp);
        }
        ready = true;
    }
    const double hw = 0.5 * w;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double t = 0.5 * M_PI * gx[i];
        const double ct = std::cos(t);
        s += gw[i] * std::exp(beta * (ct - 1.0)) * std::cos(xi * hw * std::sin(t)) * ct;
    }
    return hw * 0.5 * M_PI * s;
}

// (cos, sin) of every projection angle, FP64, once per geometry: the spread
// tests every angle against every tile and the interpolation reads one per
// point
__global__ void k_angle_table(const double *__restrict__ theta, double2 *__restrict__ cs, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double s, c;
        sincos(theta[i], &s, &c);
        cs[i] = make_double2(c, s);
    }
}

int nufft2d_fine_n(int64_t N) {
    int64_t nf = 2 * N;
    while (!is_2_3_smooth(nf) || nf < 2 * NU_W) ++nf;
    return (int)nf;
}

void nufft2d_init(Ctx &c, Geom2D &g) {
    Nufft2D &u = g.nu;
    u.cs.alloc(sizeof(double2) * g.nang);
    hipLaunchKernelGGL(k_angle_table, dim3(ceil_div(g.nang, 256)), dim3(256), 0, c.stream(),
                       g.d_theta.as<double>(), u.cs.as<double2>(), (int)g.nang);
    LC_HIP(hipGetLastError());
    const int N = (int)g.N;
    const int64_t nf = nufft2d_fine_n(N);
    u.nf = (int)nf;
    u.w = NU_W;
    u.beta = 2.30 * NU_W;
    std::vector<float> corr(N);
    for (int a = 0; a < N; ++a) {
        const int m = a - N / 2;
        const double ph = es_phihat(u.beta, NU_W, (double)m * 2.0 * M_PI / (double)nf);
        const double sgn = (m & 1) ? -1.0 : 1.0;   // (-1)^m from the centred fine grid
        corr[a] = (float)(sgn / ph);
    }
    u.corr.alloc(sizeof(float) * N);
    // stream-ordered copy: a legacy-stream copy would have to wait on other
    // threads' streams, some of which may be capturing graphs
    LC_HIP(hipMemcpyAsync(u.corr.as(), corr.data(), sizeof(float) * N, hipMemcpyHostToDevice, c.stream()));
    c.sync();
    LC_REQUIRE(fft_make_plan(u.nf, u.plan_nf), "nufft: fine grid size not 2^p 3^q");
    u.det_fast = fft_make_plan((int)g.nrad, u.plan_det);
    u.ready = true;
}

namespace {
constexpr int NU_NT = 256, NU_ME = 16, NU_MEC = 32;   // rows / columns: NT*ME elements in LDS
// compiled column passes: 512 threads x NU_ME, B = a power of two columns
constexpr int NU_NT_CT = 512;
struct ColShape {
    int B, lsh;
};
ColShape col_shape(int nf, int N) {
    ColShape cs{1, 0};
    while (2 * cs.B * nf <= NU_NT_CT * NU_ME && cs.B < N) {
        cs.B *= 2;
        ++cs.lsh;
    }
    return cs;
}

SpreadGeom spread_geom(const Geom2D &g) {
    SpreadGeom s;
    s.nang = (int)g.nang;
    s.nrad = (int)g.nrad;
    s.nf = g.nu.nf;
    s.ntx = ceil_div(g.nu.nf, NU_T);
    s.dr = 2.0 * M_PI / (double)g.nrad;
    s.sc = (double)g.nu.nf / (2.0 * M_PI);
    s.beta = (float)g.nu.beta;
    return s;
}

// ns slices of g.nang sinogram rows each
void det_transform(Ctx &c, const Geom2D &g, int sign, const float *real_in, const float2 *cplx_in,
                   float2 *cplx_out, float *real_out, int ns) {
    const int n = (int)g.nrad, nang = (int)g.nang, nrows = ns * nang;
    const float2 *tw = c.twiddles<float>(n);
    const float scale = 1.0f / (float)n;
    if (g.nu.det_fast && n <= NU_NT * NU_ME) {
        const int nl = std::max(1, std::min(NU_NT * NU_ME / n, 64));
        const size_t sh = sizeof(float2) * lds_round(nl * n);
        const dim3 grid(ceil_div(nang, nl), ns);
        auto launch = [&](auto len) {
            constexpr int LC = decltype(len)::value;
            if (sign < 0)
                hipLaunchKernelGGL((k_det_fft<-1, NU_NT, NU_ME, LC>), grid, dim3(NU_NT), sh, c.stream(),
                                   real_in, cplx_in, cplx_out, real_out, nang, n, nl, g.nu.plan_det, tw, scale);
            else
                hipLaunchKernelGGL((k_det_fft<1, NU_NT, NU_ME, LC>), grid, dim3(NU_NT), sh, c.stream(),
                                   real_in, cplx_in, cplx_out, real_out, nang, n, nl, g.nu.plan_det, tw, scale);
        };
        if (!with_ct_len(n, launch)) launch(std::integral_constant<int, 0>{});
    } else {
        dim3 grid(nrows);
        if (sign < 0)
            hipLaunchKernelGGL(k_det_dft<-1>, grid, dim3(128), 0, c.stream(), real_in, cplx_in,
                               cplx_out, real_out, nrows, n, tw, scale);
        else
            hipLaunchKernelGGL(k_det_dft<1>, grid, dim3(128), 0, c.stream(), real_in, cplx_in,
                               cplx_out, real_out, nrows, n, tw, scale);
    }
    LC_HIP(hipGetLastError());
}
} // namespace

namespace {
struct NuWork {
    float2 *C, *F, *V;
    // the binned spread's keys, indices, bin runs, launch order and sort
    // scratch (null when the spread is not binned)
    unsigned *key_a = nullptr, *key_b = nullptr, *val_a = nullptr, *val_b = nullptr;
    unsigned *bstart = nullptr, *bend = nullptr, *okey_a = nullptr, *okey_b = nullptr, *oval_a = nullptr,
             *order = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_bytes = 0;
};

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// bits of sort keys 0 .. total - 1
int key_bits2(uint64_t total) {
    int b = 1;
    while (b < 32 && (1ull << b) < total) ++b;
    return b;
}

// the binned spread with bins of edge b needs nf a multiple of 2 b
// (conflict-free wrapped phases) and 32-bit point indices and keys;
// otherwise the tile gather.  Buffers are sized for 16-cell bins (the most
// bins of any shape), so every shape that fits nf shares them.
bool shape2_fits(int64_t ns, int64_t npts, int nf, int b) {
    const uint64_t nbins = (uint64_t)(nf / b) * (nf / b);
    return nf % (2 * b) == 0 && (uint64_t)(ns * npts) < (1ull << 32) && ns * nbins < (1ull << 32);
}
bool spread_binned(int64_t ns, int64_t npts, int nf) { return shape2_fits(ns, npts, nf, shape2_edge(0)); }
constexpr int B2 = 16;   // bin edge the buffers are sized for

// the binned spread's buffers beyond C, F and V: 4 x 32-bit per point, 6 x
// 32-bit per bin, and the larger of the two sorts' scratch
size_t spread2_extra_bytes(int64_t ns, int64_t npts, int nf, size_t *tmp_bytes = nullptr) {
    const uint64_t n = (uint64_t)(ns * npts), total = (uint64_t)ns * (nf / B2) * (nf / B2);
    size_t tb = 0, to = 0;
    LC_HIP(rocprim::radix_sort_pairs(nullptr, tb, (unsigned *)nullptr, (unsigned *)nullptr, (unsigned *)nullptr,
                                     (unsigned *)nullptr, (size_t)n, 0, key_bits2(total), (hipStream_t)0));
    LC_HIP(rocprim::radix_sort_pairs(nullptr, to, (unsigned *)nullptr, (unsigned *)nullptr, (unsigned *)nullptr,
                                     (unsigned *)nullptr, (size_t)total, 0, B2_ORDER_BITS, (hipStream_t)0));
    const size_t t = al256(std::max(tb, to));
    if (tmp_bytes) *tmp_bytes = t;
    return 4 * al256(sizeof(unsigned) * n) + 6 * al256(sizeof(unsigned) * total) + t;
}

NuWork nu_work(Ctx &c, const Geom2D &g, int ns, bool spread = false) {
    const int N = (int)g.N, nf = g.nu.nf;
    const int64_t npts = g.nang * g.nrad;
    const size_t bC = al256(sizeof(float2) * (size_t)ns * npts);
    const size_t bF = al256(sizeof(float2) * (size_t)ns * nf * nf);
    const size_t bV = al256(sizeof(float2) * (size_t)ns * nf * N);
    spread = spread && spread_binned(ns, npts, nf);
    size_t tmp = 0;
    const size_t extra = spread ? spread2_extra_bytes(ns, npts, nf, &tmp) : 0;
    char *ws = reinterpret_cast<char *>(c.scratch(3, bC + bF + bV + extra));
    NuWork w;
    w.C = reinterpret_cast<float2 *>(ws);
    w.F = reinterpret_cast<float2 *>(ws + bC);
    w.V = reinterpret_cast<float2 *>(ws + bC + bF);
    if (spread) {
        const size_t bK = al256(sizeof(unsigned) * (size_t)(ns * npts));
        const size_t bB = al256(sizeof(unsigned) * (size_t)ns * (nf / B2) * (nf / B2));
        char *q = ws + bC + bF + bV;
        w.key_a = reinterpret_cast<unsigned *>(q);
        w.key_b = reinterpret_cast<unsigned *>(q + bK);
        w.val_a = reinterpret_cast<unsigned *>(q + 2 * bK);
        w.val_b = reinterpret_cast<unsigned *>(q + 3 * bK);
        q += 4 * bK;
        w.bstart = reinterpret_cast<unsigned *>(q);   // bstart and bend are adjacent
        w.bend = reinterpret_cast<unsigned *>(q + bB);
        w.okey_a = reinterpret_cast<unsigned *>(q + 2 * bB);
        w.okey_b = reinterpret_cast<unsigned *>(q + 3 * bB);
        w.oval_a = reinterpret_cast<unsigned *>(q + 4 * bB);
        w.order = reinterpret_cast<unsigned *>(q + 5 * bB);
        w.sort_tmp = q + 6 * bB;
        w.sort_bytes = tmp;
    }
    return w;
}

// type-1 spread of the stack by sorted bins of shape S: C -> F (ns x nf x nf)
template <typename S> void spread2_binned(Ctx &c, const Geom2D &g, const NuWork &w, int ns) {
    const SpreadGeom sg = spread_geom(g);
    Bins2 bn;
    bn.nb = sg.nf / S::B;
    bn.nbins = (unsigned)bn.nb * bn.nb;
    bn.npts = (int)(g.nang * g.nrad);
    const int64_t n = (int64_t)ns * bn.npts;
    const unsigned total = (unsigned)ns * bn.nbins;
    const double2 *cs = g.nu.cs.as<double2>();
    hipLaunchKernelGGL(k_bin2_keys<S>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c.stream(), cs, sg, bn,
                       w.key_a, w.val_a, n);
    LC_HIP(hipGetLastError());
    size_t tb = w.sort_bytes;
    LC_HIP(rocprim::radix_sort_pairs(w.sort_tmp, tb, w.key_a, w.key_b, w.val_a, w.val_b, (size_t)n, 0,
                                     key_bits2(total), c.stream()));
    hipLaunchKernelGGL(k_zero_u32_2, dim3(1024), dim3(256), 0, c.stream(), w.bstart,
                       (int64_t)(w.bend - w.bstart) + total);
    hipLaunchKernelGGL(k_bin2_ranges, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c.stream(), w.key_b, n,
                       w.bstart, w.bend);
    hipLaunchKernelGGL(k_bin2_order_keys, dim3(ceil_div(total, 256u)), dim3(256), 0, c.stream(), w.bstart,
                       w.bend, bn, w.okey_a, w.oval_a, total);
    LC_HIP(hipGetLastError());
    tb = w.sort_bytes;
    LC_HIP(rocprim::radix_sort_pairs(w.sort_tmp, tb, w.okey_a, w.okey_b, w.oval_a, w.order, (size_t)total, 0,
                                     B2_ORDER_BITS, c.stream()));
    hipLaunchKernelGGL(k_zero_f2, dim3(2048), dim3(256), 0, c.stream(), w.F, (int64_t)ns * sg.nf * sg.nf);
    LC_HIP(hipGetLastError());
    const size_t sh = sizeof(float2) * S::RC;
    int64_t first = 0;   // the phase's first entry in the launch order
    for (int ph = 0; ph < 4; ++ph) {
        const int py = ph >> 1, px = ph & 1;
        const int64_t nblk = (int64_t)ns * ((bn.nb - py + 1) / 2) * ((bn.nb - px + 1) / 2);
        if (nblk == 0) continue;
        hipLaunchKernelGGL(k_spread2d_bin_heavy<S>, dim3((unsigned)std::min<int64_t>(nblk, 256)),
                           dim3(64 * S::NWH), sh * S::NWH, c.stream(), w.C, w.F, cs, w.val_b, w.bstart, w.bend,
                           w.order + first, (unsigned)nblk, sg, bn);
        hipLaunchKernelGGL(k_spread2d_bin<S>, dim3((unsigned)nblk), dim3(64), sh, c.stream(), w.C, w.F, cs,
                           w.val_b, w.bstart, w.bend, w.order + first, sg, bn);
        LC_HIP(hipGetLastError());
        first += nblk;
    }
}

// type 1: C (ns x nang x nrad complex) -> vol = Re(T1 C) / nrad
void type1(Ctx &c, const Geom2D &g, const NuWork &w, float *vol, int ns) {
    const Nufft2D &u = g.nu;
    const int N = (int)g.N, nf = u.nf;
    const SpreadGeom sg = spread_geom(g);
    if (w.key_a) {
        // the context's bin shape for this fine-grid size, if it fits this stack
        int shape = c.nu2_shape(nf);
        if (!shape2_fits(ns, g.nang * g.nrad, nf, shape2_edge(shape))) shape = 0;
        with_shape2(shape, [&](auto sh) { spread2_binned<decltype(sh)>(c, g, w, ns); });
    } else {
        const size_t shs = sizeof(float2) * NU_WAVES * NU_T * NU_T + sizeof(float) * NU_WAVES * 64 * 2 * NU_W;
        hipLaunchKernelGGL(k_spread2d, dim3(sg.ntx * sg.ntx, ns), dim3(64 * NU_WAVES), shs, c.stream(),
                           w.C, w.F, g.nu.cs.as<double2>(), sg);
        LC_HIP(hipGetLastError());
    }
    const float2 *tw = c.twiddles<float>(nf);
    const int nl = std::min(NU_NT * NU_ME / nf, nf);
    const size_t shr = sizeof(float2) * lds_round(nl * nf);
    if (!with_ct_len(nf, [&](auto len) {
            hipLaunchKernelGGL((k_fine_rows_t1<NU_NT, NU_ME, decltype(len)::value>), dim3(ceil_div(nf, nl), ns),
                               dim3(NU_NT), shr, c.stream(), w.F, w.V, nf, N, nl, u.plan_nf, tw);
        }))
        hipLaunchKernelGGL((k_fine_rows_t1<NU_NT, NU_ME, 0>), dim3(ceil_div(nf, nl), ns), dim3(NU_NT), shr,
                           c.stream(), w.F, w.V, nf, N, nl, u.plan_nf, tw);
    LC_HIP(hipGetLastError());
    if (!with_ct_len(nf, [&](auto len) {
            const ColShape cs = col_shape(nf, N);
            hipLaunchKernelGGL((k_fine_cols_t1<NU_NT_CT, NU_ME, decltype(len)::value>), dim3(ceil_div(N, cs.B), ns),
                               dim3(NU_NT_CT), sizeof(float2) * lds_round(cs.B * nf), c.stream(), w.V, vol,
                               u.corr.as<float>(), nf, N, cs.B, cs.lsh, 1.0f / (float)g.nrad, u.plan_nf, tw);
        })) {
        const int B = std::min(NU_NT * NU_MEC / nf, N);
        hipLaunchKernelGGL((k_fine_cols_t1<NU_NT, NU_MEC, 0>), dim3(ceil_div(N, B), ns), dim3(NU_NT),
                           sizeof(float2) * lds_round(B * nf), c.stream(), w.V, vol, u.corr.as<float>(), nf, N,
                           B, 0, 1.0f / (float)g.nrad, u.plan_nf, tw);
    }
    LC_HIP(hipGetLastError());
}

// type 2: vol (ns x N x N real) -> C = T2 vol
void type2(Ctx &c, const Geom2D &g, const NuWork &w, const float *vol, int ns) {
    const Nufft2D &u = g.nu;
    const int N = (int)g.N, nf = u.nf;
    const int64_t npts = g.nang * g.nrad;
    const float2 *tw = c.twiddles<float>(nf);
    if (!with_ct_len(nf, [&](auto len) {
            const ColShape cs = col_shape(nf, N);
            hipLaunchKernelGGL((k_fine_cols_t2<NU_NT_CT, NU_ME, decltype(len)::value>), dim3(ceil_div(N, cs.B), ns),
                               dim3(NU_NT_CT), sizeof(float2) * lds_round(cs.B * nf), c.stream(), vol, w.V,
                               u.corr.as<float>(), nf, N, cs.B, cs.lsh, u.plan_nf, tw);
        })) {
        const int B = std::min(NU_NT * NU_MEC / nf, N);
        hipLaunchKernelGGL((k_fine_cols_t2<NU_NT, NU_MEC, 0>), dim3(ceil_div(N, B), ns), dim3(NU_NT),
                           sizeof(float2) * lds_round(B * nf), c.stream(), vol, w.V, u.corr.as<float>(), nf,
                           N, B, 0, u.plan_nf, tw);
    }
    LC_HIP(hipGetLastError());
    const int nl = std::min(NU_NT * NU_ME / nf, nf);
    const size_t shr = sizeof(float2) * lds_round(nl * nf);
    if (!with_ct_len(nf, [&](auto len) {
            hipLaunchKernelGGL((k_fine_rows_t2<NU_NT, NU_ME, decltype(len)::value>), dim3(ceil_div(nf, nl), ns),
                               dim3(NU_NT), shr, c.stream(), w.V, w.F, nf, N, nl, u.plan_nf, tw);
        }))
        hipLaunchKernelGGL((k_fine_rows_t2<NU_NT, NU_ME, 0>), dim3(ceil_div(nf, nl), ns), dim3(NU_NT), shr,
                           c.stream(), w.V, w.F, nf, N, nl, u.plan_nf, tw);
    LC_HIP(hipGetLastError());
    const SpreadGeom sg = spread_geom(g);
    hipLaunchKernelGGL(k_interp2d, dim3(ceil_div(npts, 256), ns), dim3(256), 0, c.stream(), w.F, w.C,
                       g.nu.cs.as<double2>(), sg, (int)npts);
    LC_HIP(hipGetLastError());
}

void nu_check(Ctx &c, Geom2D &g, int ns) {
    if (!g.nu.ready) nufft2d_init(c, g);
    LC_REQUIRE(g.nu.nf <= NU_NT * NU_ME, "nufft: fine grid too large (image size above 2048)");
    LC_REQUIRE(ns >= 1 && ns <= 65535, "nufft: 1 to 65535 slices per call");
}
} // namespace

// backproj2d: sino (ns x nang x nrad real, device) -> vol (ns x N x N real, device)
void backproj2d_dev(Ctx &c, Geom2D &g, const float *sino, float *vol, int ns) {
    nu_check(c, g, ns);
    const NuWork w = nu_work(c, g, ns, true);
    det_transform(c, g, -1, sino, nullptr, w.C, nullptr, ns);
    type1(c, g, w, vol, ns);
}

// forward2d: vol (ns x N x N real, device) -> sino (ns x nang x nrad real, device)
void forward2d_dev(Ctx &c, Geom2D &g, const float *vol, float *sino, int ns) {
    nu_check(c, g, ns);
    const NuWork w = nu_work(c, g, ns);
    type2(c, g, w, vol, ns);
    det_transform(c, g, 1, nullptr, w.C, nullptr, sino, ns);
}

// sysmat2d(x, grid) = Re(T1(T2 x)) / nrad (src/gradient.cpp:59-69)
void sysmat2d_nufft_dev(Ctx &c, Geom2D &g, const float *x, float *y, int ns) {
    nu_check(c, g, ns);
    const NuWork w = nu_work(c, g, ns, true);
    type2(c, g, w, x, ns);
    type1(c, g, w, y, ns);
}

// Times the back-projection of nslices zero sinograms on g's grid with every
// compiled bin shape of the spread that fits its fine grid, installs the
// fastest in c.nu2_tuning for that fine-grid size and returns its index
// (best_ms: one back-projection).  Grids the binned spread does not serve
// keep shape 0 (0 ms).
int nufft2d_autotune(Ctx &c, Geom2D &g, int nslices, double *best_ms) {
    nu_check(c, g, nslices);
    const int nf = g.nu.nf;
    const int64_t npts = g.nang * g.nrad;
    if (best_ms) *best_ms = 0.0;
    if (!spread_binned(nslices, npts, nf)) return 0;
    DBuf sino(sizeof(float) * (size_t)(nslices * npts)), vol(sizeof(float) * (size_t)nslices * g.N * g.N);
    LC_HIP(hipMemsetAsync(sino.as(), 0, sizeof(float) * (size_t)(nslices * npts), c.stream()));
    const auto saved = c.nu2_tuning;
    hipEvent_t e0, e1;
    LC_HIP(hipEventCreate(&e0));
    LC_HIP(hipEventCreate(&e1));
    int best = 0;
    float tbest = 1e30f;
    try {
        for (int i = 0; i < N_SHAPES2; ++i) {
            if (!shape2_fits(nslices, npts, nf, shape2_edge(i))) continue;
            c.nu2_tuning[nf] = i;
            backproj2d_dev(c, g, sino.as<float>(), vol.as<float>(), nslices);   // warm-up
            LC_HIP(hipEventRecord(e0, c.stream()));
            for (int r = 0; r < 3; ++r) backproj2d_dev(c, g, sino.as<float>(), vol.as<float>(), nslices);
            LC_HIP(hipEventRecord(e1, c.stream()));
            LC_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            LC_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms / 3 < tbest) {
                tbest = ms / 3;
                best = i;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c.nu2_tuning = saved;
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c.nu2_tuning = saved;
    c.nu2_tuning[nf] = best;
    if (best_ms) *best_ms = tbest;
    return best;
}

int nufft2d_num_shapes() { return N_SHAPES2; }

// device bytes of a 2-D back-projection's binned spread beyond C, F and V
// (0 when the stack spreads by tile gather)
size_t nufft2d_spread_bytes(int64_t ns, int64_t npts, int nf) {
    return spread_binned(ns, npts, nf) ? spread2_extra_bytes(ns, npts, nf) : 0;
}

} // namespace lc

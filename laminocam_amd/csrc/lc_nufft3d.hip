// 3-D laminography NUFFT projectors on gfx950: forward, backproj (with
// optional centre-of-rotation shifts) and the NUFFT normal operator
// (src/projection.cpp:42-131, src/gradient.cpp:32-41).
//
//   forward(x)  = Re(ifftshift2(ifft2(fftshift2(w . T2 x)))) / (nrows ncols)
//   backproj(y) = Re(T1(w . e^{-i(qx dx + qy dy)} . ifftshift2(fft2(fftshift2 y)))) / (nrows ncols)
//   sysmat(x)   = Re(T1(T2 x)) / (nrows ncols)                (no mask)
// T1/T2: type-1/type-2 NUFFT on the grid R_i (qX_k, qY_j, 0), modes {n1,n2,n3}
// with m_x <-> n3, m_y <-> n2, m_z <-> n1 (include/nufft.h:37-58); w = 0 where
// any coordinate exceeds pi (include/polar_grid.h:80-86).
//
// Method as in 2-D: exponential-of-semicircle kernel, width 8, fine grid 2n
// per axis (2^p 3^q, at least 16), centred so points map to cells without a
// shift; the (-1)^m phase and 1/psi_hat sit in per-axis correction factors.
//   type 1: deterministic sorted-bin spreading -> fine grid, then one pruned
//           FFT pass per axis (x, y, z), deconvolution in each pass;
//   type 2: the reverse passes, then 8x8x8 interpolation per point.
// Detector transforms are line FFTs along x then y with the reference's shift
// index maps; widths that are not 2^p 3^q use Bluestein's algorithm with an
// FP64-built chirp and chirp spectrum.
#include <cmath>
#include <complex>

#include <rocprim/device/device_radix_sort.hpp>

#include "lc_engine.hpp"

namespace lc {

extern __shared__ __align__(16) char lc_smem[];

namespace {

constexpr int W3 = 8;                       // kernel width (cells)
constexpr int LN_NT = 256, LN_ME = 32;      // line passes: NT * ME complex in LDS

// The compiled line passes hold NT * LN_ME_CT values per block (32 KiB), so
// a CU holds five blocks.
constexpr int LN_ME_CT = 16;
template <typename F> bool with_line_len(int L, F &&f) { return with_ct_len(L, f); }

// exp(beta (sqrt(1 - z^2) - 1)) for |z| < 1, else 0.  The hardware square
// root (within 1 ulp; s lies in (0, 1]) instead of the correctly rounded
// sequence: one instruction instead of about fifteen, a relative weight
// error near 1e-6, inside the kernel's own approximation error
__device__ __forceinline__ float es3(float z, float beta) {
    const float s = 1.0f - z * z;
    return s > 0.0f ? __expf(beta * (__builtin_amdgcn_sqrtf(s) - 1.0f)) : 0.0f;
}

// value of v in lane `lane` (the same for every lane: a scalar broadcast, no LDS)
__device__ __forceinline__ float rl_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double rl_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int wrapi(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}

struct Grid3 {
    int nproj, nrows, ncols;
    int nf[3];          // fine grid (z, y, x)
    double dX, dY;      // lattice steps 2 pi / ncols, 2 pi / nrows
    float beta;
    int mask;           // 1: drop points with any |coordinate| > pi
};

__device__ __forceinline__ void point3(const double *rot, int p, int j, int k, const Grid3 &g,
                                       double &x, double &y, double &z) {
    const double *r = rot + 6 * (int64_t)p;
    const double qx = ((double)k + 0.5) * g.dX - M_PI;
    const double qy = ((double)j + 0.5) * g.dY - M_PI;
    x = qx * r[0] + qy * r[3];
    y = qx * r[1] + qy * r[4];
    z = qx * r[2] + qy * r[5];
}

__device__ __forceinline__ bool inside_pi(double x, double y, double z) {
    return fabs(x) <= M_PI && fabs(y) <= M_PI && fabs(z) <= M_PI;
}

// ---------------------------------------------------------------------------
// type-2 interpolation: a wave per point.  Each lane first finds the footprint
// of one of the wave's 64 consecutive points; the wave then takes the points
// in turn with its lanes over the 8 x 8 (y, x) footprint, walking the 8 z
// planes (each step reads 8 rows of 8 contiguous fine-grid values), and sums
// the 64 lanes in a fixed order.  Lane i writes point i, so the stores are
// contiguous.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_interp3d(const float2 *__restrict__ F, float2 *__restrict__ C, const double *__restrict__ rot,
           Grid3 g) {
    static_assert(W3 == 8, "the lane layout assumes a footprint of 8 per axis");
    const int64_t npts = (int64_t)g.nproj * g.nrows * g.ncols;
    const int lane = threadIdx.x & 63;
    const int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
    if (base >= npts) return;   // whole wave
    const int64_t q = base + lane;
    const double hw = 0.5 * W3;
    const float ihw = 1.0f / (float)hw;
    double gz = 0.0, gy = 0.0, gx = 0.0;
    int z0 = 0, y0 = 0, x0 = 0, ok = 0;
    if (q < npts) {
        const int k = (int)(q % g.ncols);
        const int64_t t = q / g.ncols;
        const int j = (int)(t % g.nrows), p = (int)(t / g.nrows);
        double x, y, z;
        point3(rot, p, j, k, g, x, y, z);
        ok = (!g.mask || inside_pi(x, y, z)) ? 1 : 0;
        gz = 0.5 * g.nf[0] + z * (g.nf[0] / (2.0 * M_PI));
        gy = 0.5 * g.nf[1] + y * (g.nf[1] / (2.0 * M_PI));
        gx = 0.5 * g.nf[2] + x * (g.nf[2] / (2.0 * M_PI));
        z0 = (int)ceil(gz - hw);
        y0 = (int)ceil(gy - hw);
        x0 = (int)ceil(gx - hw);
    }
    const int iy = lane >> 3, ix = lane & 7;
    const int nlast = (int)(npts - 1 - base < 63 ? npts - 1 - base : 63);
    float2 res = make_float2(0.f, 0.f);
    for (int j = 0; j <= nlast; ++j) {
        if (!__builtin_amdgcn_readlane(ok, j)) continue;   // masked point: 0
        const double gzj = rl_f64(gz, j), gyj = rl_f64(gy, j), gxj = rl_f64(gx, j);
        const int z0j = __builtin_amdgcn_readlane(z0, j);
        const int y0j = __builtin_amdgcn_readlane(y0, j);
        const int x0j = __builtin_amdgcn_readlane(x0, j);
        // lane (iy, ix) holds the y weight of row iy, the x weight of column
        // ix, and the z weight of plane ix (read by every lane in the z walk)
        const float wzl = es3((float)((double)(z0j + ix) - gzj) * ihw, g.beta);
        const float wyl = es3((float)((double)(y0j + iy) - gyj) * ihw, g.beta);
        const float wxl = es3((float)((double)(x0j + ix) - gxj) * ihw, g.beta);
        const int yy = wrapi(y0j + iy, g.nf[1]), xx = wrapi(x0j + ix, g.nf[2]);
        float rr = 0.f, ri = 0.f;
#pragma unroll
        for (int a = 0; a < W3; ++a) {
            const float wza = rl_f32(wzl, a);
            const int zz = wrapi(z0j + a, g.nf[0]);
            const float2 v = F[((int64_t)zz * g.nf[1] + yy) * g.nf[2] + xx];
            rr += wza * v.x;
            ri += wza * v.y;
        }
        const float w = wyl * wxl;
        float sr = w * rr, si = w * ri;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, 64);
            si += __shfl_xor(si, off, 64);
        }
        if (lane == j) res = make_float2(sr, si);
    }
    if (q < npts) C[q] = res;
}

// include/fft.h:53-74 (phase_shift2d): detector spectrum C[p][j][k] *=
// exp(-i (qx_k dx_p + qy_j dy_p)), qx_k = (k + 1/2) 2 pi / ncols - pi,
// qy_j = (j + 1/2) 2 pi / nrows - pi; once per point, before the spread
__global__ void k_phase_shift3d(float2 *__restrict__ C, const double *__restrict__ shifts, Grid3 g,
                                int64_t np) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int64_t pj = i / g.ncols;
    const int kk = (int)(i - pj * g.ncols);
    const int pp = (int)(pj / g.nrows), jj = (int)(pj - (int64_t)pp * g.nrows);
    const double qx = ((double)kk + 0.5) * g.dX - M_PI;
    const double qy = ((double)jj + 0.5) * g.dY - M_PI;
    double s, c;
    sincos(qx * shifts[2 * pp] + qy * shifts[2 * pp + 1], &s, &c);
    const float cr = (float)c, si = (float)-s;
    const float2 cv = C[i];
    C[i] = make_float2(cv.x * cr - cv.y * si, cv.x * si + cv.y * cr);
}

// ---------------------------------------------------------------------------
// type-1 spreading by footprint-corner bins.  The fine grid is cut into
// BZ x BY x BX bins of footprint corners (a point's corner is its first
// footprint cell, ceil(g - W3/2) per axis, taken modulo nf).  k_bin_keys
// gives every point its bin; a stable radix sort (rocPRIM) orders the points
// by bin, keeping point order inside a bin; k_bin_ranges marks each bin's
// run.  A wave then owns one bin (a bin of more than B3_HEAVY points gets
// NW3_HEAVY waves, each taking a consecutive part of the bin's points): it
// takes the points in order and spreads each one's whole 8 x 8 x 8 footprint
// (lanes over 8 x 8 in x, y; 8 steps in z) into the wave's private LDS copy
// of the bin's padded region (bin + W3 - 1 cells per axis: 11 x 11 x
// 15 cells, 15.5 KiB, so a CU holds ten bins).  Every point is spread once,
// with all 64 lanes inside its footprint, and no point is tested against a
// bin it does not belong to.  The copies are summed in a fixed order and
// added into Fp, the fine grid extended by W3 - 1 cells per axis (not
// periodic).  Bins P apart along an axis have disjoint padded regions when
// P b >= b + W3 - 1 (P = 3 for b = 4, 2 for b = 8), so the bins run in
// PZ PY PX = 18 phases (the bin's index modulo P along each axis), each adding
// without conflicts, fullest bins first; k_fold3d then wraps the overhang
// back onto the periodic grid.  Every sum runs in a fixed order: the result
// is deterministic.  On the 32 x 511^2 back-projection from 720 x 128 x 511
// points the spread takes about 14 ms (8^3 bins in 8 phases: 16 ms), against
// 53 ms for a tile-gather spread that searched every tile's candidate points.
// ---------------------------------------------------------------------------
constexpr int PXS = 16;                     // padded row stride (cells)
// bins P apart along an axis have disjoint padded regions when P b >= b + W3 - 1
constexpr int phase_period(int b) { return (b + W3 - 1 + b - 1) / b; }

// A bin shape (edges in corner cells along z, y, x) the spread is compiled
// for: its padded region, phase periods and launch phases.  Region cell
// (z, y, x) -> LDS index with x XOR-ed with 8 on every other row pair, so the
// 8 rows x 8 columns of one z step fall on distinct bank pairs.
template <int BZ_, int BY_, int BX_> struct Shape3 {
    static constexpr int BZ = BZ_, BY = BY_, BX = BX_;
    static constexpr int RZ = BZ + W3 - 1, RY = BY + W3 - 1, RX = BX + W3 - 1;
    static constexpr int PZ = phase_period(BZ), PY = phase_period(BY), PX = phase_period(BX);
    static constexpr int NPH = PZ * PY * PX;   // launch phases
    // bits of the launch-order key: the phase above a 26-bit descending count
    static constexpr int ORDER_BITS = 26 + (NPH > 16 ? 5 : NPH > 8 ? 4 : NPH > 4 ? 3 : NPH > 2 ? 2 : 1);
    static constexpr int PC = RZ * RY * PXS;    // cells of one region copy
    // waves of a heavy bin: as many region copies as fit 64 KiB
    static constexpr int NWH = 65536 / (8 * PC) < 1 ? 1 : 65536 / (8 * PC);
    static_assert(RX <= PXS, "a region row must fit the padded row stride");
    static_assert(ORDER_BITS <= 32, "launch-order key too wide");
    __device__ __forceinline__ static int tidx(int z, int y, int x) {
        return (z * RY + y) * PXS + (x ^ (((y >> 1) & 1) << 3));
    }
};
// The shapes the autotuner (spread3d_autotune) chooses from, per fine-grid
// shape; shape 0 is the default.  On the 32 x 511^2 back-projection
// (720 x 128 x 511 points) the whole call measured 30.9 ms with 4 x 4 x 8
// bins, 31.2 with 8 x 4 x 4, 31.8 with 4 x 4 x 4, 32.8 with 2 x 4 x 8 and 33.0
// with 8 x 8 x 8.
constexpr int N_SHAPES3 = 3;
template <typename F> void with_shape3(int i, F &&f) {
    switch (i) {
    case 1: f(Shape3<8, 4, 4>{}); break;
    case 2: f(Shape3<8, 8, 8>{}); break;
    default: f(Shape3<4, 4, 8>{}); break;
    }
}
constexpr int MAX_ORDER_BITS = 31;

struct Bins3 {
    int nb[3];          // bins per axis, ceil(nf / bin edge)
    int pf[3];          // padded grid dims nf + W3 - 1
};

// fine-grid position g, footprint corner l (unwrapped) of point (p, j, k)
struct Corner3 {
    double gz, gy, gx;
    int lz, ly, lx;
    bool ok;    // false: dropped by the mask
};
__device__ __forceinline__ Corner3 corner3(const double *rot, int p, int j, int k, const Grid3 &g) {
    Corner3 c;
    double x, y, z;
    point3(rot, p, j, k, g, x, y, z);
    c.ok = !g.mask || inside_pi(x, y, z);
    const double hw = 0.5 * W3;
    c.gz = 0.5 * g.nf[0] + z * (g.nf[0] / (2.0 * M_PI));
    c.gy = 0.5 * g.nf[1] + y * (g.nf[1] / (2.0 * M_PI));
    c.gx = 0.5 * g.nf[2] + x * (g.nf[2] / (2.0 * M_PI));
    c.lz = (int)ceil(c.gz - hw);
    c.ly = (int)ceil(c.gy - hw);
    c.lx = (int)ceil(c.gx - hw);
    return c;
}

// key[i]: the bin of point i's wrapped corner, or nbins (sorted last) for a
// point the mask drops; val[i] = i
template <typename S>
__global__ void k_bin_keys(const double *__restrict__ rot, Grid3 g, Bins3 bn, unsigned *__restrict__ key,
                           unsigned *__restrict__ val, int64_t npts, unsigned nbins) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const int64_t pj = i / g.ncols;
    const int k = (int)(i - pj * g.ncols);
    const int p = (int)(pj / g.nrows), j = (int)(pj - (int64_t)p * g.nrows);
    const Corner3 c = corner3(rot, p, j, k, g);
    unsigned b = nbins;
    if (c.ok) {
        const int bz = wrapi(c.lz, g.nf[0]) / S::BZ, by = wrapi(c.ly, g.nf[1]) / S::BY,
                  bx = wrapi(c.lx, g.nf[2]) / S::BX;
        b = (unsigned)(((int64_t)bz * bn.nb[1] + by) * bn.nb[2] + bx);
    }
    key[i] = b;
    val[i] = (unsigned)i;
}

// [start[b], end[b]) is bin b's run in the sorted keys (start = end = 0 when
// empty: both arrays cleared first)
__global__ void k_bin_ranges(const unsigned *__restrict__ skey, int64_t n, unsigned *__restrict__ start,
                             unsigned *__restrict__ end, unsigned nbins) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned b = skey[i];
    if (b >= nbins) return;
    if (i == 0 || skey[i - 1] != b) start[b] = (unsigned)i;
    if (i == n - 1 || skey[i + 1] != b) end[b] = (unsigned)(i + 1);
}

// the bins in launch order: grouped by phase (the bin's indices modulo the
// phase periods), fullest first inside a phase, so a phase's long-running
// dense bins start first and its tail is short.  key = phase << 26 |
// (2^26 - 1 - count), val = bin
template <typename S>
__global__ void k_bin_order_keys(const unsigned *__restrict__ bstart, const unsigned *__restrict__ bend,
                                 Bins3 bn, unsigned *__restrict__ key, unsigned *__restrict__ val,
                                 unsigned nbins) {
    const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    const unsigned bx = b % bn.nb[2], by = (b / bn.nb[2]) % bn.nb[1], bz = b / (bn.nb[2] * bn.nb[1]);
    const unsigned ph = ((bz % S::PZ) * S::PY + by % S::PY) * S::PX + bx % S::PX;
    const unsigned cnt = min(bend[b] - bstart[b], (1u << 26) - 1);
    key[b] = (ph << 26) | ((1u << 26) - 1 - cnt);
    val[b] = b;
}

// Bins of more than B3_HEAVY points are spread by S::NWH waves, each taking a
// consecutive part of the bin's points into its own copy of the region; the
// copies are summed in wave order.  Whether a bin is heavy depends only on
// its own count.
constexpr unsigned B3_HEAVY = 2048;

// spread bin bid's points with the block's NW waves into their region copies
// (reg: NW x RZ x RY x PXS) and add the copies' fixed-order sum into Fp
template <typename S, int NW>
__device__ __forceinline__ void spread3_bin(const float2 *__restrict__ C, float2 *__restrict__ Fp,
                                            const double *__restrict__ rot, const unsigned *__restrict__ sval,
                                            int bid, unsigned s0, unsigned s1, const Grid3 &g, const Bins3 &bn,
                                            float2 *acc) {
    constexpr int PC = S::PC, RY = S::RY, RZ = S::RZ, RX = S::RX;
    const int b[3] = {bid / (bn.nb[1] * bn.nb[2]), (bid / bn.nb[2]) % bn.nb[1], bid % bn.nb[2]};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float2 *my = acc + wave * PC;
    for (int i = lane; i < PC; i += 64) my[i] = make_float2(0.f, 0.f);
    const int c0[3] = {S::BZ * b[0], S::BY * b[1], S::BX * b[2]};
    const float ihw = 1.0f / (0.5f * W3);
    const int dxl = lane & 7, dyl = lane >> 3;
    // wave w takes the w-th part of the bin's points, in order
    const unsigned per = (s1 - s0 + NW - 1) / NW;
    const unsigned ws = s0 + wave * per, we = min(s1, ws + per);
    for (unsigned base = ws; base < we; base += 64) {
        const unsigned i = base + lane;
        const bool ok = i < we;
        float2 cv = make_float2(0.f, 0.f);
        int tz = 0, ty = 0, tx = 0;
        float fz = 0.f, fy = 0.f, fx = 0.f;
        if (ok) {
            const unsigned pt = sval[i];
            const unsigned pj = pt / (unsigned)g.ncols;
            const int k = (int)(pt - pj * (unsigned)g.ncols);
            const int p = (int)(pj / (unsigned)g.nrows), j = (int)(pj - (unsigned)p * (unsigned)g.nrows);
            const Corner3 c = corner3(rot, p, j, k, g);
            // the corner taken modulo nf, relative to the bin: in [0, bin edge)
            tz = wrapi(c.lz, g.nf[0]) - c0[0];
            ty = wrapi(c.ly, g.nf[1]) - c0[1];
            tx = wrapi(c.lx, g.nf[2]) - c0[2];
            fz = (float)((double)c.lz - c.gz);
            fy = (float)((double)c.ly - c.gy);
            fx = (float)((double)c.lx - c.gx);
            cv = C[pt];
        }
        unsigned long long pts = __ballot(ok);
        while (pts) {
            const int s2 = __ffsll((long long)pts) - 1;
            pts &= pts - 1;
            const int pz0 = __builtin_amdgcn_readlane(tz, s2);
            const int py0 = __builtin_amdgcn_readlane(ty, s2);
            const int px0 = __builtin_amdgcn_readlane(tx, s2);
            const float vr = rl_f32(cv.x, s2), vi = rl_f32(cv.y, s2);
            const float wxy = es3(((float)dxl + rl_f32(fx, s2)) * ihw, g.beta) *
                              es3(((float)dyl + rl_f32(fy, s2)) * ihw, g.beta);
            // lane dz holds wz[dz]
            const float wzl = es3(((float)dxl + rl_f32(fz, s2)) * ihw, g.beta);
            const int base_idx = S::tidx(pz0, py0 + dyl, px0 + dxl);
#pragma unroll
            for (int dz = 0; dz < W3; ++dz) {
                const float wgt = wxy * rl_f32(wzl, dz);
                float2 &a = my[base_idx + dz * RY * PXS];
                a.x += wgt * vr;
                a.y += wgt * vi;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < RZ * RY * RX; i += 64 * NW) {
        const int z = i / (RY * RX), y = (i / RX) % RY, x = i % RX;
        const int zz = c0[0] + z, yy = c0[1] + y, xx = c0[2] + x;
        if (zz >= bn.pf[0] || yy >= bn.pf[1] || xx >= bn.pf[2]) continue;
        const int li = S::tidx(z, y, x);
        float2 s = acc[li];
        for (int w = 1; w < NW; ++w) {
            s.x += acc[w * PC + li].x;
            s.y += acc[w * PC + li].y;
        }
        float2 &f = Fp[((int64_t)zz * bn.pf[1] + yy) * bn.pf[2] + xx];
        f.x += s.x;
        f.y += s.y;
    }
}

// one wave per light bin of this phase (order: the phase's bins, fullest first)
template <typename S>
__global__ void __launch_bounds__(64)
k_spread3d_bin(const float2 *__restrict__ C, float2 *__restrict__ Fp, const double *__restrict__ rot,
               const unsigned *__restrict__ sval, const unsigned *__restrict__ bstart,
               const unsigned *__restrict__ bend, const unsigned *__restrict__ order, Grid3 g, Bins3 bn) {
    const int bid = (int)order[blockIdx.x];
    const unsigned s0 = bstart[bid], s1 = bend[bid];
    if (s0 >= s1 || s1 - s0 > B3_HEAVY) return;   // empty, or a heavy bin
    spread3_bin<S, 1>(C, Fp, rot, sval, bid, s0, s1, g, bn, reinterpret_cast<float2 *>(lc_smem));
}

// the phase's heavy bins (a prefix of its order): each block takes bins
// j, j + gridDim.x, ... until the first light one
template <typename S>
__global__ void __launch_bounds__(64 * S::NWH)
k_spread3d_bin_heavy(const float2 *__restrict__ C, float2 *__restrict__ Fp, const double *__restrict__ rot,
                     const unsigned *__restrict__ sval, const unsigned *__restrict__ bstart,
                     const unsigned *__restrict__ bend, const unsigned *__restrict__ order, unsigned nphase,
                     Grid3 g, Bins3 bn) {
    for (unsigned j = blockIdx.x; j < nphase; j += gridDim.x) {
        const int bid = (int)order[j];
        const unsigned s0 = bstart[bid], s1 = bend[bid];
        if (s1 - s0 <= B3_HEAVY) break;   // the order is fullest first: the rest are light
        spread3_bin<S, S::NWH>(C, Fp, rot, sval, bid, s0, s1, g, bn, reinterpret_cast<float2 *>(lc_smem));
        __syncthreads();   // the copies are cleared for the next bin only after every wave has flushed
    }
}

__global__ void k_zero_u32(unsigned *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_zero2(float2 *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_float2(0.f, 0.f);
}

// F = Fp wrapped onto the periodic fine grid: F[z][y][x] sums the (up to 8)
// cells of Fp at (z + az nf0, y + ay nf1, x + ax nf2), az, ay, ax in {0, 1},
// in that fixed order
__global__ void k_fold3d(const float2 *__restrict__ Fp, float2 *__restrict__ F, Grid3 g, Bins3 bn) {
    const int64_t n = (int64_t)g.nf[0] * g.nf[1] * g.nf[2];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i / ((int64_t)g.nf[1] * g.nf[2]));
        const int64_t rem = i - (int64_t)z * g.nf[1] * g.nf[2];
        const int y = (int)(rem / g.nf[2]), x = (int)(rem - (int64_t)y * g.nf[2]);
        float2 s = make_float2(0.f, 0.f);
        for (int az = 0; az < 2; ++az) {
            const int zz = z + az * g.nf[0];
            if (zz >= bn.pf[0]) break;
            for (int ay = 0; ay < 2; ++ay) {
                const int yy = y + ay * g.nf[1];
                if (yy >= bn.pf[1]) break;
                for (int ax = 0; ax < 2; ++ax) {
                    const int xx = x + ax * g.nf[2];
                    if (xx >= bn.pf[2]) break;
                    const float2 v = Fp[((int64_t)zz * bn.pf[1] + yy) * bn.pf[2] + xx];
                    s.x += v.x;
                    s.y += v.y;
                }
            }
        }
        F[i] = s;
    }
}

// ---------------------------------------------------------------------------
// pruned line passes between the fine grid and the modes along one axis.
// Array (outer, L, inner); row mode when inner == 1 (nl lines per block),
// otherwise column mode with B consecutive inner indices per block.
// DIR 1 (type 1): in (outer, nf, inner) -> FFT+ -> out (outer, nm, inner),
//                 out[a] = X[m_a mod nf] * corr[a] (* scale, real part if ROUT)
// DIR 2 (type 2): in (outer, nm, inner) (* corr, real if RIN) placed at
//                 m_a mod nf, zeros elsewhere -> FFT- -> out (outer, nf, inner)
// ---------------------------------------------------------------------------
// one FFT of nl lines in LDS: row layout (rowm) or nl = 1 << lsh interleaved
// lines; LC > 0: length LC compiled (unrolled stages), else the runtime plan
template <int SIGN, int NT, int MAXE, int LC>
__device__ __forceinline__ void line_fft(float2 *buf, bool rowm, int nl, int lsh, int L, const FftPlan &plan,
                                         const float2 *__restrict__ tw) {
    if constexpr (LC > 0) {
        if (rowm) fft_lds_ct<float, LC, SIGN, NT, MAXE, false>(buf, nl, 0, tw);
        else fft_lds_ct<float, LC, SIGN, NT, MAXE, true>(buf, nl, lsh, tw);
    } else {
        if (rowm) fft_lds<float, SIGN, NT, MAXE, false>(buf, nl, L, plan, tw);
        else fft_lds<float, SIGN, NT, MAXE, true>(buf, nl, 0, plan, tw);
    }
}

__device__ __forceinline__ int m2f(int a, int nm, int nf) {
    const int m = a - nm / 2;
    return m < 0 ? m + nf : m;
}

template <int DIR, bool RIN, bool ROUT, int NT, int MAXE, int LC>
__global__ void __launch_bounds__(NT)
k_nu_axis(const void *__restrict__ vin, void *__restrict__ vout, int64_t outer, int64_t inner,
          int nf, int nm, int nl, int lsh, const float *__restrict__ corr, float scale, FftPlan plan,
          const float2 *__restrict__ tw) {
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    if constexpr (LC > 0) nf = LC;
    const int Lin = DIR == 1 ? nf : nm, Lout = DIR == 1 ? nm : nf;
    const bool rowm = inner == 1;
    int64_t o0, i0;
    int nloc;
    if (rowm) {
        o0 = (int64_t)blockIdx.x * nl;
        i0 = 0;
        nloc = (int)min((int64_t)nl, outer - o0);
    } else {
        o0 = blockIdx.y;
        i0 = (int64_t)blockIdx.x * nl;
        nloc = (int)min((int64_t)nl, inner - i0);
    }
    // element e of local line l: address of (o, e, i)
    auto addr = [&](int l, int e, int L) -> int64_t {
        return rowm ? (o0 + l) * L + e : (o0 * L + e) * inner + i0 + l;
    };
    auto lds = [&](int l, int e) -> int { return rowm ? lds_row(l, e, nf) : lds_col(l, e, nl); };
    if (DIR == 2) {
        for (int idx = threadIdx.x; idx < nl * nf; idx += NT)
            buf[rowm ? swz(idx) : swz(idx)] = make_float2(0.f, 0.f);
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nl * Lin; idx += NT) {
        int l, e;
        if (rowm) { l = idx / Lin; e = idx - l * Lin; }
        else { e = idx / nl; l = idx - e * nl; }
        if (l >= nloc) continue;
        const int64_t q = addr(l, e, Lin);
        float2 v = RIN ? make_float2(reinterpret_cast<const float *>(vin)[q], 0.f)
                       : reinterpret_cast<const float2 *>(vin)[q];
        if (DIR == 1) {
            buf[lds(l, e)] = v;
        } else {
            const float c = corr[e];
            buf[lds(l, m2f(e, nm, nf))] = make_float2(v.x * c, v.y * c);
        }
    }
    if (DIR == 1) {
        // lines beyond nloc hold stale LDS: clear them so the FFT reads zeros
        for (int idx = threadIdx.x; idx < nl * nf; idx += NT) {
            int l, e;
            if (rowm) { l = idx / nf; e = idx - l * nf; }
            else { e = idx / nl; l = idx - e * nl; }
            if (l >= nloc) buf[lds(l, e)] = make_float2(0.f, 0.f);
        }
    }
    __syncthreads();
    line_fft<DIR == 1 ? 1 : -1, NT, MAXE, LC>(buf, rowm, nl, lsh, nf, plan, tw);
    for (int idx = threadIdx.x; idx < nl * Lout; idx += NT) {
        int l, e;
        if (rowm) { l = idx / Lout; e = idx - l * Lout; }
        else { e = idx / nl; l = idx - e * nl; }
        if (l >= nloc) continue;
        const int64_t q = addr(l, e, Lout);
        if (DIR == 1) {
            const float2 v = buf[lds(l, m2f(e, nm, nf))];
            const float c = corr[e] * scale;
            if (ROUT) reinterpret_cast<float *>(vout)[q] = v.x * c;
            else reinterpret_cast<float2 *>(vout)[q] = make_float2(v.x * c, v.y * c);
        } else {
            reinterpret_cast<float2 *>(vout)[q] = buf[lds(l, e)];
        }
    }
}

// ---------------------------------------------------------------------------
// detector line transform along one axis with the reference shift maps:
// buf[k] = in[(k + din) mod n]; X = DFT_sign(buf); out[k] = X[(k + dout) mod n]
// Stockham when n is 2^p 3^q, else Bluestein of length Lb:
// X_m = a_m sum_k (x_k a_k) b_{m-k},  a_k = e^{s i pi k^2/n},  b_k = conj(a_k)
// computed as a_m * FFT+(FFT-(x a) * bhat) with bhat = FFT-(b) / Lb.
// ---------------------------------------------------------------------------
template <bool RIN, bool ROUT, int NT, int MAXE, int LC>
__global__ void __launch_bounds__(NT)
k_det_axis(const void *__restrict__ vin, void *__restrict__ vout, int64_t outer, int64_t inner,
           int n, int nl, int lsh, int din, int dout, int sign, float scale, int blue, FftPlan plan,
           const float2 *__restrict__ tw, const float2 *__restrict__ chirp,
           const float2 *__restrict__ bhat) {
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    const int L = LC > 0 ? LC : blue ? plan.L : n;
    const bool rowm = inner == 1;
    int64_t o0, i0;
    int nloc;
    if (rowm) {
        o0 = (int64_t)blockIdx.x * nl;
        i0 = 0;
        nloc = (int)min((int64_t)nl, outer - o0);
    } else {
        o0 = blockIdx.y;
        i0 = (int64_t)blockIdx.x * nl;
        nloc = (int)min((int64_t)nl, inner - i0);
    }
    auto addr = [&](int l, int e) -> int64_t {
        return rowm ? (o0 + l) * n + e : (o0 * n + e) * inner + i0 + l;
    };
    auto lds = [&](int l, int e) -> int { return rowm ? lds_row(l, e, L) : lds_col(l, e, nl); };
    for (int idx = threadIdx.x; idx < nl * L; idx += NT) {
        int l, e;
        if (rowm) { l = idx / L; e = idx - l * L; }
        else { e = idx / nl; l = idx - e * nl; }
        float2 v = make_float2(0.f, 0.f);
        if (l < nloc && e < n) {
            int src = e + din;
            if (src >= n) src -= n;
            const int64_t q = addr(l, src);
            v = RIN ? make_float2(reinterpret_cast<const float *>(vin)[q], 0.f)
                    : reinterpret_cast<const float2 *>(vin)[q];
            if (blue) {
                float2 a = chirp[e];
                if (sign < 0) a.y = -a.y;
                v = cmul(v, a);
            }
        }
        buf[lds(l, e)] = v;
    }
    __syncthreads();
    if (!blue) {
        if (sign < 0) line_fft<-1, NT, MAXE, LC>(buf, rowm, nl, lsh, L, plan, tw);
        else line_fft<1, NT, MAXE, LC>(buf, rowm, nl, lsh, L, plan, tw);
    } else {
        line_fft<-1, NT, MAXE, LC>(buf, rowm, nl, lsh, L, plan, tw);
        const float2 *bh = bhat + (sign < 0 ? 0 : L);
        for (int idx = threadIdx.x; idx < nl * L; idx += NT) {
            int l, e;
            if (rowm) { l = idx / L; e = idx - l * L; }
            else { e = idx / nl; l = idx - e * nl; }
            float2 *p = buf + lds(l, e);
            *p = cmul(*p, bh[e]);
        }
        __syncthreads();
        line_fft<1, NT, MAXE, LC>(buf, rowm, nl, lsh, L, plan, tw);
    }
    for (int idx = threadIdx.x; idx < nl * n; idx += NT) {
        int l, e;
        if (rowm) { l = idx / n; e = idx - l * n; }
        else { e = idx / nl; l = idx - e * nl; }
        if (l >= nloc) continue;
        int m = e + dout;
        if (m >= n) m -= n;
        float2 v = buf[lds(l, m)];
        if (blue) {
            float2 a = chirp[m];
            if (sign < 0) a.y = -a.y;
            v = cmul(v, a);
        }
        const int64_t q = addr(l, e);
        if (ROUT) reinterpret_cast<float *>(vout)[q] = v.x * scale;
        else reinterpret_cast<float2 *>(vout)[q] = make_float2(v.x * scale, v.y * scale);
    }
}

bool smooth23(int64_t n) { return is_2_3_smooth(n); }

int fine_size(int64_t n) {
    int64_t f = std::max<int64_t>(2 * n, 2 * W3);
    while (!smooth23(f)) ++f;
    return (int)f;
}

// psi_hat of the width-8 kernel by 96-point Gauss-Legendre quadrature (FP64)
double es_phihat3(double beta, double xi) {
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
    const double hw = 0.5 * W3;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double t = 0.5 * M_PI * gx[i];
        const double ct = std::cos(t);
        s += gw[i] * std::exp(beta * (ct - 1.0)) * std::cos(xi * hw * std::sin(t)) * ct;
    }
    return hw * 0.5 * M_PI * s;
}

} // namespace

int nufft3d_fine_n(int64_t n) { return fine_size(n); }

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void det_line_init(Ctx &c, DetLine &d, int n) {
    d.n = n;
    if (n <= LN_NT * LN_ME && fft_make_plan(n, d.plan)) {
        d.blue = false;
        return;
    }
    // any L >= 2n - 1 gives the exact DFT: the smallest 2^p 3^q, or a length
    // with a compiled line FFT when one is within 25% of it
    int64_t L = 2 * (int64_t)n - 1;
    while (!is_2_3_smooth(L)) ++L;
    for (int cl = (int)(2 * (int64_t)n - 1); cl <= L + L / 4 && cl <= 2048; ++cl)
        if (with_line_len(cl, [](auto) {})) {
            L = cl;
            break;
        }
    LC_REQUIRE(L <= LN_NT * LN_ME, "detector transform: width too large");
    LC_REQUIRE(fft_make_plan((int)L, d.plan), "detector transform: no Bluestein length");
    d.blue = true;
    // chirp a_k = e^{+i pi k^2 / n} (the kernel conjugates it for sign -1)
    std::vector<float2> ch(n);
    std::vector<std::complex<double>> a(n);
    for (int k = 0; k < n; ++k) {
        const int64_t q = ((int64_t)k * k) % (2 * (int64_t)n);
        const double ang = M_PI * (double)q / (double)n;
        a[k] = std::complex<double>(std::cos(ang), std::sin(ang));
        ch[k] = make_float2((float)a[k].real(), (float)a[k].imag());
    }
    // bhat for sign s: b_k = conj(a_k^s) = e^{-s i pi k^2/n} on k and L-k
    // (L >= 2n - 1, so the two runs do not meet); bhat = FFT-(b) / L in FP64,
    // [sign -1 | sign +1].  b_k = b_{L-k} pairs e^{-2 pi i k m / L} with its
    // conjugate: bhat_m = (b_0 + sum_{0<k<n} b_k 2 cos(2 pi k m / L)) / L, a
    // real combination of the b_k, so the sign +1 table is the conjugate of
    // the sign -1 one.  cos(2 pi q / L) comes from a table over q = k m mod L.
    std::vector<double> cq(L);
    for (int64_t q = 0; q < L; ++q) cq[q] = std::cos(2.0 * M_PI * (double)q / (double)L);
    std::vector<float2> bh(2 * L);
    for (int64_t m = 0; m < L; ++m) {
        std::complex<double> acc = a[0];
        int64_t q = 0;
        for (int k = 1; k < n; ++k) {
            q += m;
            if (q >= L) q -= L;
            acc += a[k] * (2.0 * cq[q]);
        }
        acc /= (double)L;
        bh[m] = make_float2((float)acc.real(), (float)acc.imag());
        bh[L + m] = make_float2((float)acc.real(), (float)-acc.imag());
    }
    d.chirp.alloc(sizeof(float2) * n);
    d.bhat.alloc(sizeof(float2) * 2 * L);
    LC_HIP(hipMemcpyAsync(d.chirp.as(), ch.data(), sizeof(float2) * n, hipMemcpyHostToDevice, c.stream()));
    LC_HIP(hipMemcpyAsync(d.bhat.as(), bh.data(), sizeof(float2) * 2 * L, hipMemcpyHostToDevice, c.stream()));
    c.sync();
}

void nufft3d_init(Ctx &c, Geom3D &g) {
    Nufft3D &u = g.nu;
    const int64_t nm[3] = {g.n1, g.n2, g.n3};
    const double beta = 2.30 * W3;
    u.beta = beta;
    std::vector<float> corr;
    for (int a = 0; a < 3; ++a) {
        u.nf[a] = fine_size(nm[a]);
        LC_REQUIRE(u.nf[a] <= LN_NT * LN_ME, "3-D NUFFT: fine grid too large");
        LC_REQUIRE(fft_make_plan(u.nf[a], u.plan[a]), "3-D NUFFT: fine size not 2^p 3^q");
        u.coff[a] = (int)corr.size();
        for (int64_t i = 0; i < nm[a]; ++i) {
            const int64_t m = i - nm[a] / 2;
            const double ph = es_phihat3(beta, (double)m * 2.0 * M_PI / (double)u.nf[a]);
            corr.push_back((float)(((m & 1) ? -1.0 : 1.0) / ph));
        }
    }
    u.corr.alloc(sizeof(float) * corr.size());
    LC_HIP(hipMemcpyAsync(u.corr.as(), corr.data(), sizeof(float) * corr.size(), hipMemcpyHostToDevice,
                          c.stream()));
    c.sync();
    det_line_init(c, u.det_x, (int)g.ncols);
    det_line_init(c, u.det_y, (int)g.nrows);
    u.ready = true;
}

namespace {

Grid3 grid3(const Geom3D &g, bool mask) {
    Grid3 r;
    r.nproj = (int)g.nproj;
    r.nrows = (int)g.nrows;
    r.ncols = (int)g.ncols;
    for (int a = 0; a < 3; ++a) r.nf[a] = g.nu.nf[a];
    r.dX = 2.0 * M_PI / (double)g.ncols;
    r.dY = 2.0 * M_PI / (double)g.nrows;
    r.beta = (float)g.nu.beta;
    r.mask = mask ? 1 : 0;
    return r;
}

struct Work3 {
    float2 *C, *F, *V1, *V2;
    // type 1 only: the spread's padded fine grid, the points' bin keys and
    // indices (unsorted, sorted), the bins' runs and the sort's scratch
    float2 *Fp;
    unsigned *key_a, *key_b, *val_a, *val_b, *bstart, *bend;
    unsigned *okey_a, *okey_b, *oval_a, *order;   // the bins' launch order
    void *sort_tmp;
    size_t sort_bytes;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int bins_of(int nf, int b) { return (nf + b - 1) / b; }
template <typename S> unsigned nbins_of(const int nf[3]) {
    return (unsigned)bins_of(nf[0], S::BZ) * bins_of(nf[1], S::BY) * bins_of(nf[2], S::BX);
}
// the most bins any compiled shape makes of this grid (work buffers fit every shape)
unsigned nbins_max(const int nf[3]) {
    unsigned m = 0;
    for (int i = 0; i < N_SHAPES3; ++i) with_shape3(i, [&](auto sh) { m = std::max(m, nbins_of<decltype(sh)>(nf)); });
    return m;
}

// bits of the sort key: bins 0 .. nbins - 1 and nbins for dropped points
int key_bits(unsigned nbins) {
    int b = 1;
    while (b < 32 && (1ull << b) <= nbins) ++b;
    return b;
}

size_t sort_tmp_bytes(int64_t npts, unsigned nbins) {
    size_t tb = 0, to = 0;
    LC_HIP(rocprim::radix_sort_pairs(nullptr, tb, (unsigned *)nullptr, (unsigned *)nullptr, (unsigned *)nullptr,
                                     (unsigned *)nullptr, (size_t)npts, 0, key_bits(nbins), (hipStream_t)0));
    LC_HIP(rocprim::radix_sort_pairs(nullptr, to, (unsigned *)nullptr, (unsigned *)nullptr, (unsigned *)nullptr,
                                     (unsigned *)nullptr, (size_t)nbins, 0, MAX_ORDER_BITS, (hipStream_t)0));
    return std::max(tb, to);
}

// bytes of the type-1 spread's buffers beyond the fine grid
size_t spread_bytes(int64_t npts, const int nf[3]) {
    const unsigned nbins = nbins_max(nf);
    return align256(sizeof(float2) * (size_t)(nf[0] + W3 - 1) * (nf[1] + W3 - 1) * (nf[2] + W3 - 1)) +
           4 * align256(sizeof(unsigned) * (size_t)npts) + 6 * align256(sizeof(unsigned) * (size_t)nbins) +
           align256(sort_tmp_bytes(npts, nbins));
}

Work3 work3(Ctx &c, const Geom3D &g, bool spread = false) {
    const Nufft3D &u = g.nu;
    const int64_t npts = (int64_t)g.nproj * g.nrows * g.ncols;
    const size_t bC = align256(sizeof(float2) * (size_t)npts);
    const size_t bF = align256(sizeof(float2) * (size_t)u.nf[0] * u.nf[1] * u.nf[2]);
    const size_t bV1 = align256(sizeof(float2) * (size_t)u.nf[0] * u.nf[1] * g.n3);
    const size_t bV2 = align256(sizeof(float2) * (size_t)u.nf[0] * g.n2 * g.n3);
    const size_t base = bC + bF + bV1 + bV2;
    Work3 w{};
    if (!spread) {
        char *ws = reinterpret_cast<char *>(c.scratch(4, base));
        w.C = reinterpret_cast<float2 *>(ws);
        w.F = reinterpret_cast<float2 *>(ws + bC);
        w.V1 = reinterpret_cast<float2 *>(ws + bC + bF);
        w.V2 = reinterpret_cast<float2 *>(ws + bC + bF + bV1);
        return w;
    }
    LC_REQUIRE(npts < (int64_t)UINT32_MAX, "backproj: more than 2^32 detector points");
    const unsigned nbins = nbins_max(u.nf);
    const size_t bFp = align256(sizeof(float2) * (size_t)(u.nf[0] + W3 - 1) * (u.nf[1] + W3 - 1) * (u.nf[2] + W3 - 1));
    const size_t bK = align256(sizeof(unsigned) * (size_t)npts);
    const size_t bB = align256(sizeof(unsigned) * (size_t)nbins);
    const size_t bT = sort_tmp_bytes(npts, nbins);
    char *ws = reinterpret_cast<char *>(c.scratch(4, base + bFp + 4 * bK + 6 * bB + align256(bT)));
    w.C = reinterpret_cast<float2 *>(ws);
    w.F = reinterpret_cast<float2 *>(ws + bC);
    w.V1 = reinterpret_cast<float2 *>(ws + bC + bF);
    w.V2 = reinterpret_cast<float2 *>(ws + bC + bF + bV1);
    char *q = ws + base;
    w.Fp = reinterpret_cast<float2 *>(q);
    q += bFp;
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
    q += 6 * bB;
    w.sort_tmp = q;
    w.sort_bytes = bT;
    return w;
}

// lines per block and grid of a line pass over (outer, L, inner): row mode
// (inner == 1) takes up to 64 whole rows; column mode up to 64 consecutive
// inner indices per block (a power of two for the compiled passes)
struct LineShape {
    int nl, lsh;
    dim3 grid;
};
LineShape line_shape(int64_t outer, int64_t inner, int L, int cap, bool pow2, const char *what) {
    LineShape sh;
    sh.lsh = 0;
    const int most = std::max(1, std::min(cap / L, 64));
    if (inner == 1) {
        sh.nl = most;
        sh.grid = dim3((unsigned)ceil_div(outer, sh.nl));
    } else {
        LC_REQUIRE(outer <= 65535, std::string(what) + ": too many planes");
        if (pow2) {
            while ((2 << sh.lsh) <= most && (1 << sh.lsh) < inner) ++sh.lsh;
            sh.nl = 1 << sh.lsh;
        } else {
            sh.nl = (int)std::max<int64_t>(1, std::min<int64_t>(most, inner));
        }
        sh.grid = dim3((unsigned)ceil_div(inner, sh.nl), (unsigned)outer);
    }
    return sh;
}

// one axis pass of the fine-grid transform
template <int DIR, bool RIN, bool ROUT>
void nu_axis(Ctx &c, const void *in, void *out, int64_t outer, int64_t inner, int nf, int nm,
             const float *corr, float scale, const FftPlan &plan) {
    const float2 *tw = c.twiddles<float>(nf);
    const bool ct = with_line_len(nf, [](auto) {});
    // runtime plans: the smaller block (fewer registers) while a line fits it
    const bool small = nf <= LN_NT * LN_ME_CT;
    const int cap = LN_NT * (small ? LN_ME_CT : LN_ME);
    LC_REQUIRE(nf <= cap, "3-D NUFFT: line too long");
    const LineShape ls = line_shape(outer, inner, nf, cap, ct, "3-D NUFFT");
    const size_t sh = sizeof(float2) * lds_round(ls.nl * nf);
    if (!with_line_len(nf, [&](auto len) {
            hipLaunchKernelGGL((k_nu_axis<DIR, RIN, ROUT, LN_NT, LN_ME_CT, decltype(len)::value>), ls.grid,
                               dim3(LN_NT), sh, c.stream(), in, out, outer, inner, nf, nm, ls.nl, ls.lsh, corr,
                               scale, plan, tw);
        })) {
        if (small)
            hipLaunchKernelGGL((k_nu_axis<DIR, RIN, ROUT, LN_NT, LN_ME_CT, 0>), ls.grid, dim3(LN_NT), sh,
                               c.stream(), in, out, outer, inner, nf, nm, ls.nl, ls.lsh, corr, scale, plan, tw);
        else
            hipLaunchKernelGGL((k_nu_axis<DIR, RIN, ROUT, LN_NT, LN_ME, 0>), ls.grid, dim3(LN_NT), sh,
                               c.stream(), in, out, outer, inner, nf, nm, ls.nl, ls.lsh, corr, scale, plan, tw);
    }
    LC_HIP(hipGetLastError());
}

// detector transform along one axis
template <bool RIN, bool ROUT>
void det_axis(Ctx &c, const DetLine &d, const void *in, void *out, int64_t outer, int64_t inner,
              int din, int dout, int sign, float scale) {
    const int L = d.blue ? d.plan.L : d.n;
    const float2 *tw = c.twiddles<float>(d.plan.L);
    const bool ct = with_line_len(L, [](auto) {});
    const bool small = L <= LN_NT * LN_ME_CT;
    const int cap = LN_NT * (small ? LN_ME_CT : LN_ME);
    LC_REQUIRE(L <= cap, "detector transform: line too long");
    const LineShape ls = line_shape(outer, inner, L, cap, ct, "detector transform");
    const size_t sh = sizeof(float2) * lds_round(ls.nl * L);
    const float2 *ch = d.blue ? d.chirp.as<float2>() : nullptr, *bh = d.blue ? d.bhat.as<float2>() : nullptr;
    if (!with_line_len(L, [&](auto len) {
            hipLaunchKernelGGL((k_det_axis<RIN, ROUT, LN_NT, LN_ME_CT, decltype(len)::value>), ls.grid,
                               dim3(LN_NT), sh, c.stream(), in, out, outer, inner, d.n, ls.nl, ls.lsh, din,
                               dout, sign, scale, d.blue ? 1 : 0, d.plan, tw, ch, bh);
        })) {
        if (small)
            hipLaunchKernelGGL((k_det_axis<RIN, ROUT, LN_NT, LN_ME_CT, 0>), ls.grid, dim3(LN_NT), sh, c.stream(),
                               in, out, outer, inner, d.n, ls.nl, ls.lsh, din, dout, sign, scale, d.blue ? 1 : 0,
                               d.plan, tw, ch, bh);
        else
            hipLaunchKernelGGL((k_det_axis<RIN, ROUT, LN_NT, LN_ME, 0>), ls.grid, dim3(LN_NT), sh, c.stream(), in,
                               out, outer, inner, d.n, ls.nl, ls.lsh, din, dout, sign, scale, d.blue ? 1 : 0,
                               d.plan, tw, ch, bh);
    }
    LC_HIP(hipGetLastError());
}

void type2_3d(Ctx &c, const Geom3D &g, const Work3 &w, const float *vol, bool mask) {
    const Nufft3D &u = g.nu;
    const int n1 = (int)g.n1, n2 = (int)g.n2, n3 = (int)g.n3;
    const float *corr = g.nu.corr.as<float>();
    // z: (n1, n2*n3) real -> (nf0, n2*n3)
    nu_axis<2, true, false>(c, vol, w.V2, 1, (int64_t)n2 * n3, u.nf[0], n1, corr + u.coff[0], 1.0f, u.plan[0]);
    // y: (nf0, n2, n3) -> (nf0, nf1, n3)
    nu_axis<2, false, false>(c, w.V2, w.V1, u.nf[0], n3, u.nf[1], n2, corr + u.coff[1], 1.0f, u.plan[1]);
    // x: rows (nf0*nf1, n3) -> (nf0*nf1, nf2)
    nu_axis<2, false, false>(c, w.V1, w.F, (int64_t)u.nf[0] * u.nf[1], 1, u.nf[2], n3,
                             corr + u.coff[2], 1.0f, u.plan[2]);
    const Grid3 gg = grid3(g, mask);
    const int64_t npts = g.nproj * g.nrows * g.ncols;
    // a wave of 64 points, four waves per block
    hipLaunchKernelGGL(k_interp3d, dim3(ceil_div(ceil_div(npts, (int64_t)64), (int64_t)4)), dim3(256), 0,
                       c.stream(), w.F, w.C, g.d_rot.as<double>(), gg);
    LC_HIP(hipGetLastError());
}

// the sorted-bin spread with bin shape S: C -> Fp (padded, not yet folded)
template <typename S> Bins3 spread3d(Ctx &c, const Geom3D &g, const Work3 &w, const Grid3 &gg) {
    const Nufft3D &u = g.nu;
    Bins3 bn;
    int64_t npad = 1;
    const int edge[3] = {S::BZ, S::BY, S::BX};
    for (int a = 0; a < 3; ++a) {
        bn.nb[a] = bins_of(u.nf[a], edge[a]);
        bn.pf[a] = u.nf[a] + W3 - 1;
        npad *= bn.pf[a];
    }
    const unsigned nbins = (unsigned)bn.nb[0] * bn.nb[1] * bn.nb[2];
    const int64_t npts = (int64_t)g.nproj * g.nrows * g.ncols;
    // points ordered by bin, point order kept inside a bin
    hipLaunchKernelGGL(k_bin_keys<S>, dim3((unsigned)ceil_div(npts, 256)), dim3(256), 0, c.stream(),
                       g.d_rot.as<double>(), gg, bn, w.key_a, w.val_a, npts, nbins);
    LC_HIP(hipGetLastError());
    size_t tb = w.sort_bytes;
    LC_HIP(rocprim::radix_sort_pairs(w.sort_tmp, tb, w.key_a, w.key_b, w.val_a, w.val_b, (size_t)npts, 0,
                                     key_bits(nbins), c.stream()));
    hipLaunchKernelGGL(k_zero_u32, dim3(1024), dim3(256), 0, c.stream(), w.bstart,
                       (int64_t)(w.bend - w.bstart) + nbins);
    hipLaunchKernelGGL(k_bin_ranges, dim3((unsigned)ceil_div(npts, 256)), dim3(256), 0, c.stream(), w.key_b,
                       npts, w.bstart, w.bend, nbins);
    hipLaunchKernelGGL(k_bin_order_keys<S>, dim3(ceil_div(nbins, 256)), dim3(256), 0, c.stream(), w.bstart,
                       w.bend, bn, w.okey_a, w.oval_a, nbins);
    LC_HIP(hipGetLastError());
    tb = w.sort_bytes;
    LC_HIP(rocprim::radix_sort_pairs(w.sort_tmp, tb, w.okey_a, w.okey_b, w.oval_a, w.order, (size_t)nbins, 0,
                                     S::ORDER_BITS, c.stream()));
    hipLaunchKernelGGL(k_zero2, dim3(2048), dim3(256), 0, c.stream(), w.Fp, npad);
    LC_HIP(hipGetLastError());
    const size_t shs = sizeof(float2) * S::PC;
    int64_t first = 0;   // the phase's first entry in the launch order
    for (int ph = 0; ph < S::NPH; ++ph) {
        // phase (rz, ry, rx): the bins whose index along each axis is r mod P
        const int r[3] = {ph / (S::PY * S::PX), (ph / S::PX) % S::PY, ph % S::PX}, per[3] = {S::PZ, S::PY, S::PX};
        int nblk = 1;
        for (int a = 0; a < 3; ++a) nblk *= std::max(0, (bn.nb[a] - r[a] + per[a] - 1) / per[a]);
        if (nblk == 0) continue;
        hipLaunchKernelGGL(k_spread3d_bin_heavy<S>, dim3(std::min(nblk, 256)), dim3(64 * S::NWH), shs * S::NWH,
                           c.stream(), w.C, w.Fp, g.d_rot.as<double>(), w.val_b, w.bstart, w.bend, w.order + first,
                           (unsigned)nblk, gg, bn);
        hipLaunchKernelGGL(k_spread3d_bin<S>, dim3(nblk), dim3(64), shs, c.stream(), w.C, w.Fp,
                           g.d_rot.as<double>(), w.val_b, w.bstart, w.bend, w.order + first, gg, bn);
        LC_HIP(hipGetLastError());
        first += nblk;
    }
    return bn;
}

void type1_3d(Ctx &c, const Geom3D &g, const Work3 &w, float *vol, bool mask, const double *shifts,
              float scale) {
    const Nufft3D &u = g.nu;
    const int n1 = (int)g.n1, n2 = (int)g.n2, n3 = (int)g.n3;
    const Grid3 gg = grid3(g, mask);
    if (shifts) {
        const int64_t np = (int64_t)g.nproj * g.nrows * g.ncols;
        hipLaunchKernelGGL(k_phase_shift3d, dim3(ceil_div(np, 256)), dim3(256), 0, c.stream(), w.C,
                           shifts, gg, np);
        LC_HIP(hipGetLastError());
    }
    LC_REQUIRE(w.Fp != nullptr, "type1_3d: no spread buffers in the work area");
    Bins3 bn;
    with_shape3(c.nu3_shape(u.nf), [&](auto sh) { bn = spread3d<decltype(sh)>(c, g, w, gg); });
    const int64_t nfine = (int64_t)u.nf[0] * u.nf[1] * u.nf[2];
    hipLaunchKernelGGL(k_fold3d, dim3((unsigned)std::min<int64_t>(8192, ceil_div(nfine, 256))), dim3(256), 0,
                       c.stream(), w.Fp, w.F, gg, bn);
    LC_HIP(hipGetLastError());
    const float *corr = g.nu.corr.as<float>();
    // x: rows (nf0*nf1, nf2) -> (nf0*nf1, n3)
    nu_axis<1, false, false>(c, w.F, w.V1, (int64_t)u.nf[0] * u.nf[1], 1, u.nf[2], n3,
                             corr + u.coff[2], 1.0f, u.plan[2]);
    // y: (nf0, nf1, n3) -> (nf0, n2, n3)
    nu_axis<1, false, false>(c, w.V1, w.V2, u.nf[0], n3, u.nf[1], n2, corr + u.coff[1], 1.0f, u.plan[1]);
    // z: (nf0, n2*n3) -> (n1, n2*n3) real, scaled
    nu_axis<1, false, true>(c, w.V2, vol, 1, (int64_t)n2 * n3, u.nf[0], n1, corr + u.coff[0], scale,
                            u.plan[0]);
}

void nu3_check(Ctx &c, Geom3D &g) {
    if (!g.nu.ready) nufft3d_init(c, g);
}

} // namespace

// forward (src/projection.cpp:42-64): vol {n1, N, N} -> proj {nproj, nrows, ncols}
void forward3d_dev(Ctx &c, Geom3D &g, const float *vol, float *proj) {
    nu3_check(c, g);
    const Work3 w = work3(c, g);
    type2_3d(c, g, w, vol, true);
    const int nr = (int)g.nrows, nc = (int)g.ncols;
    // fftshift2 -> ifft2 -> ifftshift2 along x then y; real part / (nrows ncols)
    det_axis<false, false>(c, g.nu.det_x, w.C, w.C, g.nproj * nr, 1, nc / 2, nc / 2 + (nc & 1), 1, 1.0f);
    det_axis<false, true>(c, g.nu.det_y, w.C, proj, g.nproj, nc, nr / 2, nr / 2 + (nr & 1), 1,
                          1.0f / ((float)nr * (float)nc));
}

// backproj (src/projection.cpp:71-131): proj -> vol; shifts (dx, dy) per
// projection or nullptr
void backproj3d_dev(Ctx &c, Geom3D &g, const float *proj, float *vol, const double *shifts) {
    nu3_check(c, g);
    const Work3 w = work3(c, g, true);
    const int nr = (int)g.nrows, nc = (int)g.ncols;
    // fftshift2 -> fft2 -> ifftshift2 along x then y
    det_axis<true, false>(c, g.nu.det_x, proj, w.C, g.nproj * nr, 1, nc / 2, nc / 2 + (nc & 1), -1, 1.0f);
    det_axis<false, false>(c, g.nu.det_y, w.C, w.C, g.nproj, nc, nr / 2, nr / 2 + (nr & 1), -1, 1.0f);
    type1_3d(c, g, w, vol, true, shifts, 1.0f / ((float)nr * (float)nc));
}

// sysmat(x, grid) (src/gradient.cpp:32-41): Re(T1(T2 x)) / (nrows ncols), no mask
void sysmat3d_nufft_dev(Ctx &c, Geom3D &g, const float *x, float *y) {
    nu3_check(c, g);
    const Work3 w = work3(c, g, true);
    type2_3d(c, g, w, x, false);
    type1_3d(c, g, w, y, false, nullptr, 1.0f / ((float)g.nrows * (float)g.ncols));
}

// device bytes of a 3-D back-projection's spread beyond the fine grid (the
// padded grid, the sort of the points by bin and its scratch)
size_t nufft3d_spread_bytes(int64_t npts, const int nf[3]) { return spread_bytes(npts, nf); }

int nufft3d_num_shapes() { return N_SHAPES3; }

// Times the back-projection of g's grid with every compiled bin shape of the
// spread (zero projections: the spread's work depends on where the points
// fall, not on their values), installs the fastest in c.nu3_tuning for g's
// fine-grid shape and returns its index (best_ms: one back-projection).
int nufft3d_autotune(Ctx &c, Geom3D &g, double *best_ms) {
    nu3_check(c, g);
    const int64_t np = g.nproj * g.nrows * g.ncols, nv = g.n1 * g.n2 * g.n3;
    DBuf proj(sizeof(float) * (size_t)np), vol(sizeof(float) * (size_t)nv);
    LC_HIP(hipMemsetAsync(proj.as(), 0, sizeof(float) * (size_t)np, c.stream()));
    const std::array<int, 3> key{g.nu.nf[0], g.nu.nf[1], g.nu.nf[2]};
    const auto saved = c.nu3_tuning;
    hipEvent_t e0, e1;
    LC_HIP(hipEventCreate(&e0));
    LC_HIP(hipEventCreate(&e1));
    int best = 0;
    float tbest = 1e30f;
    try {
        for (int i = 0; i < N_SHAPES3; ++i) {
            c.nu3_tuning[key] = i;
            backproj3d_dev(c, g, proj.as<float>(), vol.as<float>(), nullptr);   // warm-up
            LC_HIP(hipEventRecord(e0, c.stream()));
            for (int r = 0; r < 2; ++r) backproj3d_dev(c, g, proj.as<float>(), vol.as<float>(), nullptr);
            LC_HIP(hipEventRecord(e1, c.stream()));
            LC_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            LC_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms / 2 < tbest) {
                tbest = ms / 2;
                best = i;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c.nu3_tuning = saved;
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c.nu3_tuning = saved;
    c.nu3_tuning[key] = best;
    if (best_ms) *best_ms = tbest;
    return best;
}

} // namespace lc

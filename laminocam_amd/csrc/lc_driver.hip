// Whole-solve drivers, device resident:
//   prepare_sinogram (include/mbir2d_prep.h:44-65): stack the datasets,
//     divide by the global maximum, {n_angles, n_rows, n_cols} ->
//     {n_rows, n_angles, n_cols};
//   the 2-D slice-stack solve of cpu_mbir2d (src/mbir2d.cpp:66-100);
//   the 3-D laminography solve MBIR (src/mbir.cpp:34-161): stack, divide by
//     the maximum, zero-pad the detector by the pad factor to odd sizes,
//     back-project with the per-projection shifts, solve on the padded volume,
//     crop back to the requested dimensions.
// The data crosses PCIe once each way; normalisation, padding, transposes and
// crops are single HBM passes (one wave per output row, coalesced along the
// contiguous axis).
#include <algorithm>
#include <cmath>
#include <vector>

#include "laminocam.h"
#include "lc_engine.hpp"

using namespace lc;

namespace {

constexpr int RED_T = 256;     // threads per block of the max reduction
constexpr int RED_MAXB = 1024; // partial maxima (blocks) of its first pass

__device__ __forceinline__ float dmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return fmax(a, b); }

// x / m rounded once, as the host's scalar division
__device__ __forceinline__ float ddiv(float x, float m) { return __fdiv_rn(x, m); }
__device__ __forceinline__ double ddiv(double x, double m) { return __ddiv_rn(x, m); }

template <typename T> __device__ T block_max(T v) {
    __shared__ T red[RED_T / 64];
    for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < RED_T / 64; ++i) v = dmax(v, red[i]);
    }
    return v;
}

// pass 1: maximum of a grid-stride share of x per block
template <typename T> __global__ __launch_bounds__(RED_T) void k_max_part(const T *x, int64_t n, T *part) {
    T v = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * RED_T + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_T)
        v = dmax(v, x[i]);
    v = block_max(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// pass 2: one block folds the partial maxima into part[RED_MAXB]
template <typename T> __global__ __launch_bounds__(RED_T) void k_max_fold(T *part, int nb) {
    T v = -INFINITY;
    for (int i = threadIdx.x; i < nb; i += RED_T) v = dmax(v, part[i]);
    v = block_max(v);
    if (threadIdx.x == 0) part[RED_MAXB] = v;
}

// out[r][a][c] = in[a][r][c] / m (m > 0; otherwise copied unchanged).
// One wave per output row (r, a).
template <typename T>
__global__ __launch_bounds__(256) void k_norm_transpose(const T *in, T *out, const T *mp, int64_t nang,
                                                        int64_t nrows, int64_t ncols) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows * nang) return;
    const int64_t r = row / nang, a = row - r * nang;
    const T m = *mp;
    const T *src = in + (a * nrows + r) * ncols;
    T *dst = out + row * ncols;
    for (int64_t c = threadIdx.x & 63; c < ncols; c += 64) {
        const T v = src[c];
        dst[c] = m > T(0) ? ddiv(v, m) : v;
    }
}

// symmetric zero padding of every projection {nrows, ncols} into
// {n2, n3} at offsets (d2, d3), after division by the maximum (m > 0), and
// rounded to the projector's FP32.  One wave per output row.
template <typename T>
__global__ __launch_bounds__(256) void k_norm_pad(const T *in, float *out, const T *mp, int64_t nproj,
                                                  int64_t nrows, int64_t ncols, int64_t n2, int64_t n3,
                                                  int64_t d2, int64_t d3) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nproj * n2) return;
    const int64_t p = row / n2, jj = row - p * n2, j = jj - d2;
    const T m = *mp;
    float *dst = out + row * n3;
    const bool inrow = j >= 0 && j < nrows;
    const T *src = in + (p * nrows + (inrow ? j : 0)) * ncols;
    for (int64_t kk = threadIdx.x & 63; kk < n3; kk += 64) {
        const int64_t k = kk - d3;
        float v = 0.0f;
        if (inrow && k >= 0 && k < ncols) {
            const T s = src[k];
            v = (float)(m > T(0) ? ddiv(s, m) : s);
        }
        dst[kk] = v;
    }
}

// out {o1, o2, o3} = in[i + d1][j + d2][k + d3]; one wave per output row
template <typename T>
__global__ __launch_bounds__(256) void k_crop3(const T *in, T *out, int64_t o1, int64_t o2, int64_t o3,
                                               int64_t i2, int64_t i3, int64_t d1, int64_t d2, int64_t d3) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= o1 * o2) return;
    const int64_t i = row / o2, j = row - i * o2;
    const T *src = in + ((i + d1) * i2 + (j + d2)) * i3 + d3;
    T *dst = out + row * o3;
    for (int64_t k = threadIdx.x & 63; k < o3; k += 64) dst[k] = src[k];
}

template <typename S, typename D> __global__ void k_convert(const S *in, D *out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        out[i] = (D)in[i];
}

int rows_grid(int64_t rows) { return ceil_div(rows, 4); }
int flat_grid(int64_t n) { return (int)std::min<int64_t>(ceil_div(n, 256), 1 << 16); }

// maximum of n elements of x into part[RED_MAXB] (device); part holds
// RED_MAXB + 1 elements
template <typename T> void device_max(Ctx &c, const T *x, int64_t n, T *part) {
    const int nb = (int)std::min<int64_t>(RED_MAXB, ceil_div(n, RED_T * 4));
    hipLaunchKernelGGL(k_max_part<T>, dim3(nb), dim3(RED_T), 0, c.stream(), x, n, part);
    hipLaunchKernelGGL(k_max_fold<T>, dim3(1), dim3(RED_T), 0, c.stream(), part, nb);
    LC_HIP(hipGetLastError());
}

// the reference's padded size: size_t(n * factor), made odd by -1 when even
// (src/mbir.cpp:41-45, computed in the element type; include/padding.h:19-22,
// computed in float)
template <typename F> int64_t pad_size(int64_t n, F factor) {
    int64_t m = (int64_t)((F)n * factor);
    if (m % 2 == 0) m -= 1;
    return m;
}

// stacked device copy of the datasets' projections {ntot, nrows, ncols}
template <typename T>
T *stack_datasets(Ctx &c, DBuf &buf, int32_t nds, const T *const *projs, const int64_t *nang,
                  int64_t nrows, int64_t ncols, bool dev, int64_t *ntot_out) {
    int64_t ntot = 0;
    for (int32_t d = 0; d < nds; ++d) {
        LC_REQUIRE(nang[d] >= 1, "driver: every dataset needs at least one projection");
        LC_REQUIRE(projs[d], "driver: NULL dataset");
        ntot += nang[d];
    }
    const size_t per = (size_t)nrows * (size_t)ncols;
    buf.alloc(sizeof(T) * per * (size_t)ntot);
    int64_t off = 0;
    for (int32_t d = 0; d < nds; ++d) {
        copy_in(c, buf.as<T>() + off * per, projs[d], sizeof(T) * per * (size_t)nang[d], dev);
        off += nang[d];
    }
    *ntot_out = ntot;
    return buf.as<T>();
}

void check_params(const lc_recon_params *p, const char *fn) {
    LC_REQUIRE(p, std::string(fn) + ": NULL params");
    LC_REQUIRE(p->regularizer == LC_REG_SPLIT_BREGMAN || p->regularizer == LC_REG_UNCONSTRAINED,
               std::string(fn) + ": unknown regularizer");
    LC_REQUIRE(p->recon_dims[0] >= 1 && p->recon_dims[1] >= 1 && p->recon_dims[2] >= 1,
               std::string(fn) + ": recon_dims must be positive");
    LC_REQUIRE(p->max_iters >= 0 && p->inner_iters >= 0, std::string(fn) + ": negative iteration count");
}

void fill_rep(Batch2D &b, bool sb, lc_report *rep) {
    if (!rep) return;
    const CgState s = b.cg_state()[0];
    rep->iters = s.iters;
    rep->outer = 0;
    rep->res = s.res;
    rep->dx = s.dx;
    rep->stop = s.stop;
    rep->pad = 0;
    if (sb) {
        const SbState q = b.sb_state()[0];
        rep->outer = q.outer;
        rep->dx = q.norm_diff;
    }
}

// CG or Split-Bregman from x0 = 0 on the batch's single problem
void solve(Batch2D &b, const lc_recon_params *p, uint32_t flags) {
    const bool graph = flags & LC_GRAPH;
    if (p->regularizer == LC_REG_UNCONSTRAINED)
        b.cg({(int)p->max_iters}, (double)p->tol, (double)p->xtol, graph, !graph, true);
    else
        b.split_bregman({p->lambda}, {p->mu}, (int)p->max_iters, (int)p->inner_iters, (double)p->tol,
                        (double)p->xtol, (flags & LC_TV_ZERO_PAD) ? 1 : 0, (flags & LC_TV_ANISO) != 0,
                        graph, !graph, true);
}

// the projector's FP32 output buffer for the right-hand side: the batch's Y
// buffer itself for an FP32 batch, else a staging buffer (see finish_rhs)
float *rhs_target(Batch2D &b, DBuf &stage) {
    if (!b.f64()) return (float *)b.Yv();
    stage.alloc(sizeof(float) * (size_t)b.size(0));
    return stage.as<float>();
}

// FP64 batch: widen the staged FP32 right-hand side into Y
void finish_rhs(Ctx &c, Batch2D &b, DBuf &stage) {
    if (!b.f64()) return;
    const int64_t n = b.size(0);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(flat_grid(n)), dim3(256), 0, c.stream(),
                       stage.as<float>(), (double *)b.Yv(), n);
    LC_HIP(hipGetLastError());
    c.sync();
    stage.release();
}

} // namespace

namespace lc {

// ---------------------------------------------------------------------------
template <typename T>
void prepare_sinogram(Ctx &c, int32_t nds, const T *const *projs, const int64_t *nang, int64_t nrows,
                      int64_t ncols, T *sino, T *max_out, uint32_t flags) {
    LC_REQUIRE(projs && nang && sino, "lc_prepare_sinogram: NULL argument");
    LC_REQUIRE(nds >= 1 && nrows >= 1 && ncols >= 1, "lc_prepare_sinogram: bad sizes");
    std::lock_guard<std::mutex> lk(c.mu);
    c.set_current();
    const bool dev = flags & LC_DEVICE_PTRS;
    DBuf raw, part, out;
    int64_t ntot = 0;
    const T *x = stack_datasets<T>(c, raw, nds, projs, nang, nrows, ncols, dev, &ntot);
    const int64_t n = ntot * nrows * ncols;
    part.alloc(sizeof(T) * (RED_MAXB + 1));
    device_max<T>(c, x, n, part.as<T>());
    T *dst = sino;
    if (!dev) {
        out.alloc(sizeof(T) * (size_t)n);
        dst = out.as<T>();
    }
    hipLaunchKernelGGL(k_norm_transpose<T>, dim3(rows_grid(nrows * ntot)), dim3(256), 0, c.stream(), x, dst,
                       part.as<T>() + RED_MAXB, ntot, nrows, ncols);
    LC_HIP(hipGetLastError());
    if (!dev) copy_out(c, sino, dst, sizeof(T) * (size_t)n, false);
    if (max_out) copy_out(c, max_out, part.as<T>() + RED_MAXB, sizeof(T), dev);
    c.sync();
}

template <typename T>
void mbir2d(Ctx &c, const T *sino, int64_t nslices, int64_t nang, int64_t ncols, const double *theta,
            const lc_recon_params *p, T *recon, lc_report *rep, uint32_t flags) {
    LC_REQUIRE(sino && theta && recon, "lc_mbir2d: NULL argument");
    check_params(p, "lc_mbir2d");
    LC_REQUIRE(nslices >= 1 && nang >= 1 && ncols >= 1, "lc_mbir2d: bad sinogram sizes");
    const int64_t n1 = p->recon_dims[0], N = p->recon_dims[2];
    LC_REQUIRE(p->recon_dims[1] == N, "lc_mbir2d: recon_dims must be square in-plane (n2 == n3)");
    LC_REQUIRE(n1 <= nslices, "lc_mbir2d: recon_dims[0] exceeds the sinogram's slice count");
    std::lock_guard<std::mutex> lk(c.mu);
    c.set_current();
    const bool dev = flags & LC_DEVICE_PTRS, f64 = sizeof(T) == 8;
    Geom2D g;
    geom2d_init(c, g, theta, nang, ncols, N);
    geom2d_build_psf(c, g, f64);
    const int64_t ns_elems = n1 * nang * ncols, nv = n1 * N * N;
    // the first n1 slices of the sinogram, as FP32 for the projector
    DBuf s32, rhs32, stage;
    s32.alloc(sizeof(float) * (size_t)ns_elems);
    if (f64) {
        stage.alloc(sizeof(T) * (size_t)ns_elems);
        copy_in(c, stage.as(), sino, sizeof(T) * (size_t)ns_elems, dev);
        hipLaunchKernelGGL((k_convert<T, float>), dim3(flat_grid(ns_elems)), dim3(256), 0, c.stream(),
                           stage.as<T>(), s32.as<float>(), ns_elems);
        LC_HIP(hipGetLastError());
        stage.release();
    } else {
        copy_in(c, s32.as(), sino, sizeof(float) * (size_t)ns_elems, dev);
    }
    Batch2D b(c, {ProbSpec{&g, (int)n1}}, f64);
    backproj2d_dev(c, g, s32.as<float>(), rhs_target(b, rhs32), (int)n1);
    finish_rhs(c, b, rhs32);
    s32.release();
    solve(b, p, flags);
    copy_out(c, recon, b.Xv(), sizeof(T) * (size_t)nv, dev);
    c.sync();
    fill_rep(b, p->regularizer == LC_REG_SPLIT_BREGMAN, rep);
}

template <typename T>
void mbir3d(Ctx &c, int32_t nds, const T *const *projs, const int64_t *nang, int64_t nrows, int64_t ncols,
            const double *theta, const double *gamma, const double *beta, const double *shifts,
            const lc_recon_params *p, T *recon, int64_t *padded_dims, lc_report *rep, uint32_t flags) {
    LC_REQUIRE(projs && nang && theta && gamma && beta && recon, "lc_mbir3d: NULL argument");
    check_params(p, "lc_mbir3d");
    LC_REQUIRE(nds >= 1 && nrows >= 1 && ncols >= 1, "lc_mbir3d: bad sizes");
    LC_REQUIRE(p->recon_dims[1] == p->recon_dims[2],
               "lc_mbir3d: recon_dims must be square in-plane (n2 == n3)");
    const T padfac = (T)p->pad_factor;
    int64_t od[3];
    for (int i = 0; i < 3; ++i) {
        od[i] = pad_size<T>(p->recon_dims[i], padfac);
        LC_REQUIRE(od[i] >= p->recon_dims[i],
                   "lc_mbir3d: the padded volume is smaller than recon_dims (pad factor below 1, or 1 "
                   "with an even dimension)");
    }
    int64_t n2 = nrows, n3 = ncols;
    if (!((double)(p->pad_factor - 1.0f) < 1e-6)) {
        n2 = pad_size<float>(nrows, p->pad_factor);
        n3 = pad_size<float>(ncols, p->pad_factor);
        LC_REQUIRE(n2 >= nrows && n3 >= ncols, "lc_mbir3d: pad factor shrinks the projections");
    }
    int64_t ntot = 0;
    for (int32_t d = 0; d < nds; ++d) {
        LC_REQUIRE(nang[d] >= 1, "lc_mbir3d: every dataset needs at least one projection");
        ntot += nang[d];
    }
    std::lock_guard<std::mutex> lk(c.mu);
    c.set_current();
    const bool dev = flags & LC_DEVICE_PTRS, f64 = sizeof(T) == 8;

    // per-projection tilt angles of the stacked datasets (src/mbir.cpp:88-98)
    std::vector<double> ga((size_t)ntot), be((size_t)ntot);
    for (int32_t d = 0, o = 0; d < nds; ++d)
        for (int64_t i = 0; i < nang[d]; ++i, ++o) {
            ga[o] = gamma[d];
            be[o] = beta[d];
        }

    DBuf raw, part, P, dsh, rhs32;
    const T *x = stack_datasets<T>(c, raw, nds, projs, nang, nrows, ncols, dev, &ntot);
    part.alloc(sizeof(T) * (RED_MAXB + 1));
    device_max<T>(c, x, ntot * nrows * ncols, part.as<T>());
    P.alloc(sizeof(float) * (size_t)(ntot * n2 * n3));
    hipLaunchKernelGGL(k_norm_pad<T>, dim3(rows_grid(ntot * n2)), dim3(256), 0, c.stream(), x, P.as<float>(),
                       part.as<T>() + RED_MAXB, ntot, nrows, ncols, n2, n3, (n2 - nrows) / 2,
                       (n3 - ncols) / 2);
    LC_HIP(hipGetLastError());
    c.sync();
    raw.release();

    Geom3D g;
    geom3d_init(c, g, theta, ga.data(), be.data(), ntot, n2, n3, od);
    geom3d_build_psf(c, g, f64);
    if (shifts) {
        dsh.alloc(sizeof(double) * 2 * (size_t)ntot);
        copy_in(c, dsh.as(), shifts, sizeof(double) * 2 * (size_t)ntot, dev);
    }
    Batch2D b(c, {ProbSpec{nullptr, (int)od[0], &g}}, f64);
    backproj3d_dev(c, g, P.as<float>(), rhs_target(b, rhs32), shifts ? dsh.as<double>() : nullptr);
    finish_rhs(c, b, rhs32);
    c.sync();
    P.release();
    solve(b, p, flags);

    // symmetric crop (include/padding.h crop3d, PadType::SYMMETRIC)
    const int64_t *rd = p->recon_dims;
    DBuf out;
    T *dst = recon;
    if (!dev) {
        out.alloc(sizeof(T) * (size_t)(rd[0] * rd[1] * rd[2]));
        dst = out.as<T>();
    }
    hipLaunchKernelGGL(k_crop3<T>, dim3(rows_grid(rd[0] * rd[1])), dim3(256), 0, c.stream(), (const T *)b.Xv(),
                       dst, rd[0], rd[1], rd[2], od[1], od[2], (od[0] - rd[0]) / 2, (od[1] - rd[1]) / 2,
                       (od[2] - rd[2]) / 2);
    LC_HIP(hipGetLastError());
    if (!dev) copy_out(c, recon, dst, sizeof(T) * (size_t)(rd[0] * rd[1] * rd[2]), false);
    c.sync();
    if (padded_dims) {
        const int64_t v[5] = {od[0], od[1], od[2], n2, n3};
        for (int i = 0; i < 5; ++i) padded_dims[i] = v[i];
    }
    fill_rep(b, p->regularizer == LC_REG_SPLIT_BREGMAN, rep);
}

#define LC_DRIVER_INST(T)                                                                              \
    template void prepare_sinogram<T>(Ctx &, int32_t, const T *const *, const int64_t *, int64_t,     \
                                      int64_t, T *, T *, uint32_t);                                   \
    template void mbir2d<T>(Ctx &, const T *, int64_t, int64_t, int64_t, const double *,              \
                            const lc_recon_params *, T *, lc_report *, uint32_t);                     \
    template void mbir3d<T>(Ctx &, int32_t, const T *const *, const int64_t *, int64_t, int64_t,      \
                            const double *, const double *, const double *, const double *,           \
                            const lc_recon_params *, T *, int64_t *, lc_report *, uint32_t);
LC_DRIVER_INST(float)
LC_DRIVER_INST(double)

} // namespace lc

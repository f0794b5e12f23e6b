// Batched CG and Split-Bregman engine for 2-D problems (stacks of slices).
//
// CG follows src/conjgrad.cpp:33-76 per problem:
//   r = y - A x0, p = r, rho = <r,r>
//   loop: Ap = A p; pAp = <p,Ap>; stop if |pAp| < 1e-10; alpha = rho/pAp;
//         x += alpha p; r -= alpha Ap; rho' = <r,r>; p = r + (rho'/rho) p;
//         res = ||r||/(||y||+1e-8); dx = ||alpha p||/(||x||+1e-8);
//         stop if res < tol or dx < xtol
// All inner products and norms accumulate in FP64 with a fixed reduction
// order (per-block partials, then one wave per problem), so a problem's
// iterates do not depend on which other problems share the batch.
// The p update is fused into the Toeplitz row pass of the next iteration and
// <p, Ap> into the inverse row pass.
//
// Split-Bregman follows src/bregman.cpp:36-94 per problem with the operator
// A - mu * laplacian, the finite differences of src/finitediff.cpp:26-90
// (boundary 0) or zero padded ones (boundary 1, src/gpu/finitediff.cu).
//
// Every element kernel is a template on the element type R: float (FP32 bulk
// arithmetic) or double (the FP64 path of the reference's Array<double>
// instantiations, src/conjgrad.cpp:82-85, src/bregman.cpp:101-104).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>

#include <type_traits>

#include "lc_engine.hpp"

namespace lc {

// threads per block of the CG vector update (CHUNK / (4 CG_UPD_NT) four-wide
// loads of each vector per thread)
constexpr int CG_UPD_NT = 512;

// CG iterations per replay of a batch's cached iteration graph
constexpr int CG_GRAPH_ITERS = 8;

// ----------------------------------------------------------------------------
// reductions
// ----------------------------------------------------------------------------
template <int NT> __device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; ++w) t += sh[w];
    __syncthreads();
    return t;   // valid in thread 0
}

// threads per problem of the scalar kernels (k_cg_fin*, k_sb_fin)
constexpr int FIN_NT = 256;

// one block of FIN_NT threads per problem: fixed-order sum of
// partial[stride*i + comp] over [b, e) (thread t takes i = b + t, b + t + FIN_NT,
// ... in order, then the block tree); valid in thread 0; every thread calls it
__device__ __forceinline__ double range_sum(const double *partial, int b, int e, int stride,
                                            int comp, double *sh) {
    double v = 0.0;
    for (int i = b + (int)threadIdx.x; i < e; i += FIN_NT) v += partial[(int64_t)i * stride + comp];
    return block_sum<FIN_NT>(v, sh);
}

template <typename R> __global__ void k_fill(R *p, R v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// arm CG states for a solve: active = max_iter > 0 (and the SB problem active)
__global__ void k_cg_arm(CgState *st, const SbState *sb, const int *max_iter, double tol,
                         double xtol, int nprob) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    CgState s = st[p];
    s.max_iter = max_iter ? max_iter[p] : s.max_iter;
    s.tol = tol;
    s.xtol = xtol;
    s.active = (s.max_iter > 0) && (sb == nullptr || sb[p].active);
    s.iters = 0;
    s.stop = 0;
    s.beta = 0.0;
    st[p] = s;
}

// r = y - (x0 == 0 ? 0 : Ap); p = 0; partials <r,r>, <y,y>
template <typename R, int NT>
__global__ void __launch_bounds__(NT)
k_cg_init(const ChunkRef *__restrict__ ch, const CgState *__restrict__ st,
          const R *__restrict__ y, const R *__restrict__ ap, R *__restrict__ r,
          R *__restrict__ p, R *__restrict__ x, double *__restrict__ cpart, int x0_zero) {
    const ChunkRef c = ch[blockIdx.x];
    if (!st[c.prob].active) return;
    double rr = 0.0, yy = 0.0;
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += NT) {
        const R yi = y[i];
        if (x0_zero) x[i] = (R)0;   // x0 = 0: x is not read, only cleared
        const R ri = x0_zero ? yi : yi - ap[i];
        r[i] = ri;
        p[i] = (R)0;
        rr += (double)ri * ri;
        yy += (double)yi * yi;
    }
    __shared__ double sh[NT / 64];
    const double a = block_sum<NT>(rr, sh);
    const double b = block_sum<NT>(yy, sh);
    if (threadIdx.x == 0) {
        cpart[blockIdx.x * 4 + 0] = a;
        cpart[blockIdx.x * 4 + 1] = b;
    }
}

__global__ void k_cg_fin0(const ProbDev *__restrict__ pd, CgState *st, const double *cpart) {
    const int p = blockIdx.x;
    CgState s = st[p];
    if (!s.active) return;
    __shared__ double sh[FIN_NT / 64];
    const double rr = range_sum(cpart, pd[p].chunk0, pd[p].chunk0 + pd[p].nchunk, 4, 0, sh);
    const double yy = range_sum(cpart, pd[p].chunk0, pd[p].chunk0 + pd[p].nchunk, 4, 1, sh);
    if (threadIdx.x == 0) {
        s.rho = rr;
        s.ynorm = sqrt(yy);
        s.beta = 0.0;
        st[p] = s;
    }
}

// pAp from Toeplitz slot partials (use_chunks=0) or chunk partials (1)
__global__ void k_cg_fin1(const ProbDev *__restrict__ pd, CgState *st, const double *slot_part,
                          const double *cpart, int use_chunks) {
    const int p = blockIdx.x;
    CgState s = st[p];
    if (!s.active) return;
    __shared__ double sh[FIN_NT / 64];
    const double pap = use_chunks
                           ? range_sum(cpart, pd[p].chunk0, pd[p].chunk0 + pd[p].nchunk, 4, 3, sh)
                           : range_sum(slot_part, pd[p].slot0, pd[p].slot0 + pd[p].nslot, 1, 0, sh);
    if (threadIdx.x == 0) {
        s.pAp = pap;
        if (fabs(pap) < 1.0e-10) {
            s.active = 0;
            s.stop = 3;
        } else {
            s.alpha = s.rho / pap;
        }
        st[p] = s;
    }
}

// x += alpha p; r -= alpha Ap; partials <r,r>, ||alpha p||^2, ||x||^2.
// Chunks start 4-element aligned (problems are 256-byte aligned, chunks
// CHUNK elements long): whole chunks issue all their 4-wide loads before any
// arithmetic; a problem's last, shorter chunk takes 4-wide steps and a scalar
// tail.
template <typename R, int NT>
__global__ void __launch_bounds__(NT)
k_cg_update(const ChunkRef *__restrict__ ch, const CgState *__restrict__ st,
            R *__restrict__ x, R *__restrict__ r, const R *__restrict__ p,
            const R *__restrict__ ap, double *__restrict__ cpart) {
    using V = vec4<R>;
    const ChunkRef c = ch[blockIdx.x];
    const CgState s = st[c.prob];
    if (!s.active) return;
    // alpha stays FP64: each product is rounded once to FP32, as the FP64
    // oracle's step is (rounding a float alpha first doubles the error that
    // CG then amplifies)
    const double alpha = s.alpha;
    double rr = 0.0, ss = 0.0, xx = 0.0;
    auto upd = [&](R pv, R apv, R &xv, R &rv) {
        const R step = (R)((double)pv * alpha);
        xv = xv + step;
        rv = rv - (R)((double)apv * alpha);
        rr += (double)rv * rv;
        ss += (double)step * step;
        xx += (double)xv * xv;
    };
    auto upd4 = [&](const V &pv, const V &av, V &xv, V &rv) {
        upd(pv.x, av.x, xv.x, rv.x);
        upd(pv.y, av.y, xv.y, rv.y);
        upd(pv.z, av.z, xv.z, rv.z);
        upd(pv.w, av.w, xv.w, rv.w);
    };
    V *x4 = reinterpret_cast<V *>(x + c.begin);
    V *r4 = reinterpret_cast<V *>(r + c.begin);
    const V *p4 = reinterpret_cast<const V *>(p + c.begin);
    const V *a4 = reinterpret_cast<const V *>(ap + c.begin);
    const int64_t n4 = (c.end - c.begin) >> 2;
    constexpr int U = CHUNK / 4 / NT;
    static_assert(U >= 1 && U * 4 * NT == CHUNK, "CHUNK must be a multiple of 4 NT");
    if (n4 == CHUNK / 4) {
        V pv[U], av[U], xv[U], rv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = threadIdx.x + u * NT;
            pv[u] = ld_stream(p4 + i);
            av[u] = ld_stream(a4 + i);
            xv[u] = x4[i];
            rv[u] = r4[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            upd4(pv[u], av[u], xv[u], rv[u]);
            const int i = threadIdx.x + u * NT;
            st_stream(x4 + i, xv[u]);
            st_stream(r4 + i, rv[u]);
        }
    } else {
        for (int64_t i = threadIdx.x; i < n4; i += NT) {
            const V pv = p4[i], av = a4[i];
            V xv = x4[i], rv = r4[i];
            upd4(pv, av, xv, rv);
            x4[i] = xv;
            r4[i] = rv;
        }
        for (int64_t i = c.begin + 4 * n4 + threadIdx.x; i < c.end; i += NT) {
            R xv = x[i], rv = r[i];
            upd(p[i], ap[i], xv, rv);
            x[i] = xv;
            r[i] = rv;
        }
    }
    __shared__ double sh[NT / 64];
    const double a = block_sum<NT>(rr, sh);
    const double b = block_sum<NT>(ss, sh);
    const double d = block_sum<NT>(xx, sh);
    if (threadIdx.x == 0) {
        cpart[blockIdx.x * 4 + 0] = a;
        cpart[blockIdx.x * 4 + 1] = b;
        cpart[blockIdx.x * 4 + 2] = d;
    }
}

__global__ void k_cg_fin2(const ProbDev *__restrict__ pd, CgState *st, const double *cpart) {
    const int p = blockIdx.x;
    CgState s = st[p];
    if (!s.active) return;
    const int b = pd[p].chunk0, e = pd[p].chunk0 + pd[p].nchunk;
    __shared__ double sh[FIN_NT / 64];
    const double rr = range_sum(cpart, b, e, 4, 0, sh);
    const double ss = range_sum(cpart, b, e, 4, 1, sh);
    const double xx = range_sum(cpart, b, e, 4, 2, sh);
    if (threadIdx.x == 0) {
        s.beta = rr / s.rho;
        s.rho = rr;
        s.res = sqrt(rr) / (s.ynorm + 1e-8);
        s.dx = sqrt(ss) / (sqrt(xx) + 1e-8);
        s.iters += 1;
        if (s.res < s.tol) { s.active = 0; s.stop = 1; }
        else if (s.dx < s.xtol) { s.active = 0; s.stop = 2; }
        else if (s.iters >= s.max_iter) { s.active = 0; s.stop = 0; }
        st[p] = s;
    }
}

// ----------------------------------------------------------------------------
// finite differences (per element of a problem's n1 x N x N stack)
// ----------------------------------------------------------------------------
struct Idx3 {
    int s, a, b, n1, n2, n3;
};

// Neighbours are read relative to the current element's own address: rel(u0,
// ss, rs)(ds, da, db) is u0[ds*ss + da*rs + db] (ss = n2*n3, rs = n3), so the
// offsets fold to constants and no element index is multiplied out per read.
template <typename R> struct Rel {
    const R *u0;
    int64_t ss;
    int rs;
    __device__ __forceinline__ R operator()(int ds, int da, int db) const {
        return u0[ds * ss + (int64_t)(da * rs + db)];
    }
};

__device__ __forceinline__ bool interior(const Idx3 &q, int s, int a, int b) {
    return s >= 1 && s + 1 < q.n1 && a >= 1 && a + 1 < q.n2 && b >= 1 && b + 1 < q.n3;
}

// a point that differs from q's own element only along axis c (the element
// itself is inside) is inside when its coordinate v along c is in [0, n_c)
__device__ __forceinline__ bool inside_c(const Idx3 &q, int c, int v) {
    const int n = c == 0 ? q.n1 : c == 1 ? q.n2 : q.n3;
    return v >= 0 && v < n;
}
__device__ __forceinline__ int coord(const Idx3 &q, int c) { return c == 0 ? q.s : c == 1 ? q.a : q.b; }

// gradient component c (0 slices, 1 rows, 2 cols) at the element offset
// (os, oa, ob) from q's, an offset along axis c or none; u reads relative to
// q's element
template <typename R, typename A>
__device__ __forceinline__ R grad_c(const A &u, const Idx3 &q, int c, int os, int oa, int ob,
                                    int boundary) {
    const int ds = c == 0, da = c == 1, db = c == 2;
    const int s = q.s + os, a = q.a + oa, b = q.b + ob;
    if (boundary == 0) {
        if (!interior(q, s, a, b)) return (R)0;
        return (u(os + ds, oa + da, ob + db) - u(os - ds, oa - da, ob - db)) / (R)2;
    }
    // both neighbours differ from q's element along axis c only
    const int v = coord(q, c) + (c == 0 ? os : c == 1 ? oa : ob);
    const R up = inside_c(q, c, v + 1) ? u(os + ds, oa + da, ob + db) : (R)0;
    const R um = inside_c(q, c, v - 1) ? u(os - ds, oa - da, ob - db) : (R)0;
    return (R)0.5 * (up - um);
}

// divergence of a 3-component field at q's element; g(c, ds, da, db) reads
// component c at an in-range offset from it
template <typename R, typename G>
__device__ __forceinline__ R div_gen(const G &g, const Idx3 &q, int boundary) {
    const int s = q.s, a = q.a, b = q.b;
    if (boundary == 0) {
        if (!interior(q, s, a, b)) return (R)0;
        return (g(0, 1, 0, 0) - g(0, -1, 0, 0)) / (R)2 + (g(1, 0, 1, 0) - g(1, 0, -1, 0)) / (R)2 +
               (g(2, 0, 0, 1) - g(2, 0, 0, -1)) / (R)2;
    }
    const R g0p = inside_c(q, 0, s + 1) ? g(0, 1, 0, 0) : (R)0;
    const R g0m = inside_c(q, 0, s - 1) ? g(0, -1, 0, 0) : (R)0;
    const R g1p = inside_c(q, 1, a + 1) ? g(1, 0, 1, 0) : (R)0;
    const R g1m = inside_c(q, 1, a - 1) ? g(1, 0, -1, 0) : (R)0;
    const R g2p = inside_c(q, 2, b + 1) ? g(2, 0, 0, 1) : (R)0;
    const R g2m = inside_c(q, 2, b - 1) ? g(2, 0, 0, -1) : (R)0;
    const R h = (R)0.5;
    return h * (g0p - g0m) + h * (g1p - g1m) + h * (g2p - g2m);
}

// laplacian = div(grad u) at q's element, evaluated directly from u
template <typename R, typename A>
__device__ __forceinline__ R lap_at(const A &u, const Idx3 &q, int boundary) {
    R acc = (R)0;
    if (boundary == 0) {
        if (!interior(q, q.s, q.a, q.b)) return (R)0;
        // component c: (G(i+e) - G(i-e))/2 with G zero off the interior
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int ds = c == 0, da = c == 1, db = c == 2;
            const R gp = grad_c<R>(u, q, c, ds, da, db, 0);
            const R gm = grad_c<R>(u, q, c, -ds, -da, -db, 0);
            acc += (gp - gm) / (R)2;
        }
        return acc;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int ds = c == 0, da = c == 1, db = c == 2;
        const R gp = inside_c(q, c, coord(q, c) + 1) ? grad_c<R>(u, q, c, ds, da, db, 1) : (R)0;
        const R gm = inside_c(q, c, coord(q, c) - 1) ? grad_c<R>(u, q, c, -ds, -da, -db, 1) : (R)0;
        acc += (R)0.5 * (gp - gm);
    }
    return acc;
}

// (s, a, b) of the element at global offset i of a problem (one 64-bit
// division pair per thread); step() advances it by NT elements
__device__ __forceinline__ Idx3 idx_of(const ProbDev &d, int64_t i) {
    const int64_t loc = i - d.off;
    Idx3 q;
    q.n2 = d.N;
    q.n3 = d.N;
    q.n1 = d.ns;
    const int64_t nn = (int64_t)d.N * d.N;
    q.s = (int)(loc / nn);
    const int64_t rem = loc - (int64_t)q.s * nn;
    q.a = (int)(rem / d.N);
    q.b = (int)(rem - (int64_t)q.a * d.N);
    return q;
}
__device__ __forceinline__ void step(Idx3 &q, int n) {
    q.b += n;
    while (q.b >= q.n3) {
        q.b -= q.n3;
        if (++q.a >= q.n2) {
            q.a = 0;
            ++q.s;
        }
    }
}

// Ap -= mu * lap(p); partial <p, Ap> (Split-Bregman operator, src/bregman.cpp:55)
template <typename R, int NT>
__global__ void __launch_bounds__(NT)
k_sb_lapdot(const ChunkRef *__restrict__ ch, const ProbDev *__restrict__ pd,
            const CgState *__restrict__ st, const R *__restrict__ p, R *__restrict__ ap,
            double *__restrict__ cpart, int dot) {
    const ChunkRef c = ch[blockIdx.x];
    if (st && !st[c.prob].active) return;
    const ProbDev d = pd[c.prob];
    const int64_t nn = (int64_t)d.N * d.N;
    double acc = 0.0;
    Idx3 q = idx_of(d, c.begin + threadIdx.x);
    // UE elements per thread per step (i, i + NT, ...): the stencil loads of
    // all of them are issued before any result is used, so a wave keeps UE
    // elements' loads in flight instead of one; acc sums in the same order
    constexpr int UE = 4;
    auto run = [&](auto full) {
        constexpr bool FULL = decltype(full)::value;
        for (int64_t i0 = c.begin + threadIdx.x; i0 < c.end; i0 += UE * NT) {
            R l[UE], a[UE], pv[UE];
#pragma unroll
            for (int u = 0; u < UE; ++u) {
                const int64_t i = i0 + u * NT;
                if (FULL || i < c.end) {
                    l[u] = lap_at<R>(Rel<R>{p + i, nn, d.N}, q, d.boundary);
                    a[u] = __builtin_nontemporal_load(ap + i);
                    pv[u] = p[i];
                }
                step(q, NT);
            }
#pragma unroll
            for (int u = 0; u < UE; ++u) {
                const int64_t i = i0 + u * NT;
                if (FULL || i < c.end) {
                    const R v = a[u] - l[u] * (R)d.mu;
                    __builtin_nontemporal_store(v, ap + i);
                    acc += (double)pv[u] * v;
                }
            }
        }
    };
    static_assert(CHUNK % (UE * NT) == 0, "a whole chunk is a whole number of steps");
    if (c.end - c.begin == CHUNK) run(std::true_type{});
    else run(std::false_type{});
    if (dot) {
        __shared__ double sh[NT / 64];
        const double t = block_sum<NT>(acc, sh);
        if (threadIdx.x == 0) cpart[blockIdx.x * 4 + 3] = t;
    }
}

// rhs = y0 - mu * div(d - b)   (src/bregman.cpp:58-61); d - b is formed at
// each stencil point as it is read (component stride = total), rounded
// exactly as a stored d - b would be
template <typename R, int NT>
__global__ void __launch_bounds__(NT)
k_sb_rhs(const ChunkRef *__restrict__ ch, const ProbDev *__restrict__ pd,
         const SbState *__restrict__ sb, const R *__restrict__ y0, R *__restrict__ rhs,
         const R *__restrict__ D, const R *__restrict__ Bv, int64_t total) {
    const ChunkRef c = ch[blockIdx.x];
    if (!sb[c.prob].active) return;
    const ProbDev d = pd[c.prob];
    const int64_t nn = (int64_t)d.N * d.N;
    Idx3 q = idx_of(d, c.begin + threadIdx.x);
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += NT, step(q, NT)) {
        auto dmb = [&](int k, int ds, int da, int db) {
            const int64_t o = k * total + i + ds * nn + (int64_t)(da * d.N + db);
            return D[o] - Bv[o];
        };
        const R dv = div_gen<R>(dmb, q, d.boundary);
        __builtin_nontemporal_store(__builtin_nontemporal_load(y0 + i) - dv * (R)d.mu, rhs + i);
    }
}

// isotropic (or per-component) shrink and Bregman update
// (src/bregman.cpp:64-84); partials ||x - x_old||^2, ||x_old||^2; x_old = x
template <typename R, int NT>
__global__ void __launch_bounds__(NT)
k_sb_shrink(const ChunkRef *__restrict__ ch, const ProbDev *__restrict__ pd,
            const SbState *__restrict__ sb, const R *__restrict__ x, R *__restrict__ D,
            R *__restrict__ Bv, R *__restrict__ xold, double *__restrict__ cpart,
            int64_t total, int aniso) {
    const ChunkRef c = ch[blockIdx.x];
    const SbState S = sb[c.prob];
    if (!S.active) return;
    const ProbDev d = pd[c.prob];
    const int64_t nn = (int64_t)d.N * d.N;
    const R thr = (R)(S.lambda / S.mu);
    double dn = 0.0, on = 0.0;
    Idx3 q = idx_of(d, c.begin + threadIdx.x);
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += NT, step(q, NT)) {
        const Rel<R> xr{x + i, nn, d.N};
        // streaming (nontemporal) accesses for the fields only this
        // element touches; x keeps ordinary loads (its neighbours are reread)
        R g[3], v[3], bk[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            g[k] = grad_c<R>(xr, q, k, 0, 0, 0, d.boundary);
            bk[k] = __builtin_nontemporal_load(Bv + k * total + i);
            v[k] = g[k] + bk[k];
        }
        if (!aniso) {
            const R sk = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const R sh = fmax((R)0, sk - thr);
            for (int k = 0; k < 3; ++k) {
                const R dk = sh * (v[k] / (sk + (R)1e-8));
                __builtin_nontemporal_store(dk, D + k * total + i);
                __builtin_nontemporal_store(bk[k] + (g[k] - dk), Bv + k * total + i);
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                const R sh = fmax((R)0, fabs(v[k]) - thr);
                const R dk = copysign(sh, v[k]);
                __builtin_nontemporal_store(dk, D + k * total + i);
                __builtin_nontemporal_store(bk[k] + (g[k] - dk), Bv + k * total + i);
            }
        }
        const R xi = x[i], xo = __builtin_nontemporal_load(xold + i);
        dn += (double)(xi - xo) * (double)(xi - xo);
        on += (double)xo * xo;
        __builtin_nontemporal_store(xi, xold + i);
    }
    __shared__ double sh[NT / 64];
    const double a = block_sum<NT>(dn, sh);
    const double b = block_sum<NT>(on, sh);
    if (threadIdx.x == 0) {
        cpart[blockIdx.x * 4 + 0] = a;
        cpart[blockIdx.x * 4 + 1] = b;
    }
}

__global__ void k_sb_fin(const ProbDev *__restrict__ pd, SbState *sb, const double *cpart) {
    const int p = blockIdx.x;
    SbState s = sb[p];
    if (!s.active) return;
    const int b = pd[p].chunk0, e = pd[p].chunk0 + pd[p].nchunk;
    __shared__ double sh[FIN_NT / 64];
    const double dn = range_sum(cpart, b, e, 4, 0, sh);
    const double on = range_sum(cpart, b, e, 4, 1, sh);
    if (threadIdx.x == 0) {
        s.norm_diff = sqrt(dn) / (sqrt(on) + 1e-8);
        s.outer += 1;
        if (s.norm_diff < s.xtol) { s.active = 0; s.stop = 2; }
        else if (s.outer >= s.outer_max) { s.active = 0; s.stop = 0; }
        sb[p] = s;
    }
}

// ----------------------------------------------------------------------------
// standalone finite differences on one n1 x n2 x n3 volume
// ----------------------------------------------------------------------------
__device__ __forceinline__ Idx3 idx_vol(int64_t i, int n1, int n2, int n3) {
    Idx3 q;
    q.n1 = n1;
    q.n2 = n2;
    q.n3 = n3;
    q.s = (int)(i / ((int64_t)n2 * n3));
    const int64_t rem = i - (int64_t)q.s * n2 * n3;
    q.a = (int)(rem / n3);
    q.b = (int)(rem - (int64_t)q.a * n3);
    return q;
}

template <typename R>
__global__ void k_fd_grad(const R *u, R *g, int n1, int n2, int n3, int boundary) {
    const int64_t n = (int64_t)n1 * n2 * n3;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const Idx3 q = idx_vol(i, n1, n2, n3);
        const Rel<R> ur{u + i, (int64_t)n2 * n3, n3};
        for (int c = 0; c < 3; ++c) g[c * n + i] = grad_c<R>(ur, q, c, 0, 0, 0, boundary);
    }
}

template <typename R>
__global__ void k_fd_div(const R *g, R *out, int n1, int n2, int n3, int boundary) {
    const int64_t n = (int64_t)n1 * n2 * n3;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const Idx3 q = idx_vol(i, n1, n2, n3);
        auto gc = [&](int c, int ds, int da, int db) {
            return g[c * n + i + ds * ((int64_t)n2 * n3) + (int64_t)(da * n3 + db)];
        };
        out[i] = div_gen<R>(gc, q, boundary);
    }
}

template <typename R>
__global__ void k_fd_lap(const R *u, R *out, int n1, int n2, int n3, int boundary) {
    const int64_t n = (int64_t)n1 * n2 * n3;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const Idx3 q = idx_vol(i, n1, n2, n3);
        out[i] = lap_at<R>(Rel<R>{u + i, (int64_t)n2 * n3, n3}, q, boundary);
    }
}

template <typename R>
static void fd_launch_t(Ctx &c, int op, const R *in, R *out, int n1, int n2, int n3, int boundary) {
    const int64_t n = (int64_t)n1 * n2 * n3;
    const int blocks = (int)std::min<int64_t>(4096, (n + 255) / 256);
    if (op == 0)
        hipLaunchKernelGGL(k_fd_grad<R>, dim3(blocks), dim3(256), 0, c.stream(), in, out, n1, n2,
                           n3, boundary);
    else if (op == 1)
        hipLaunchKernelGGL(k_fd_div<R>, dim3(blocks), dim3(256), 0, c.stream(), in, out, n1, n2,
                           n3, boundary);
    else
        hipLaunchKernelGGL(k_fd_lap<R>, dim3(blocks), dim3(256), 0, c.stream(), in, out, n1, n2,
                           n3, boundary);
    LC_HIP(hipGetLastError());
}

void fd_launch(Ctx &c, int op, const float *in, float *out, int n1, int n2, int n3, int boundary) {
    fd_launch_t(c, op, in, out, n1, n2, n3, boundary);
}
void fd_launch(Ctx &c, int op, const double *in, double *out, int n1, int n2, int n3,
               int boundary) {
    fd_launch_t(c, op, in, out, n1, n2, n3, boundary);
}

// ----------------------------------------------------------------------------
// Batch2D
// ----------------------------------------------------------------------------
Batch2D::Batch2D(Ctx &c, const std::vector<ProbSpec> &probs, bool f64, bool class_streams)
    : c_(c), f64_(f64), probs_(probs) {
    LC_REQUIRE(!probs.empty(), "batch: no problems");
    build();
    const char *e = std::getenv("LC_CLASS_STREAMS");
    class_streams_ = class_streams && !(e && e[0] == '0') && classes_.size() > 1;
    if (class_streams_) c_.ensure_aux();
}

// call f(R{}) with R the batch's element type
#define LC_ETYPE(f) (f64_ ? (f)(double{}) : (f)(float{}))

template <typename R> void Batch2D::toep_mid(const ToepClass &k, const CgState *st, cpx<R> *U) {
    if (k.g3) toep3d_mid<R>(c_, k, st, U);
    else toep_launch_cols<R>(c_, k, dprob_.as<ProbDev>(), st, U);
}

void Batch2D::build() {
    const int np = (int)probs_.size();
    hprob_.resize(np);
    // 2-D problems: classes by N (slices of every problem of that size share
    // each launch); 3-D volumes: one class each
    std::map<int, std::vector<int>> by_n;
    std::vector<int> vols;
    for (int p = 0; p < np; ++p) {
        LC_REQUIRE(probs_[p].ns >= 1, "batch: ns must be >= 1");
        if (Geom3D *g3 = probs_[p].g3) {
            LC_REQUIRE(probs_[p].g == nullptr, "batch: problem with both a 2-D and a 3-D geometry");
            if (f64_ ? !g3->psf64_ready : !g3->psf_ready) geom3d_build_psf(c_, *g3, f64_);
            probs_[p].ns = (int)g3->n1;
            vols.push_back(p);
            continue;
        }
        Geom2D *g = probs_[p].g;
        LC_REQUIRE(g, "batch: NULL geometry");
        if (f64_ ? !g->psf64_ready : !g->psf_ready) geom2d_build_psf(c_, *g, f64_);
        by_n[(int)g->N].push_back(p);
    }
    std::vector<std::pair<int, std::vector<int>>> groups;
    for (auto &kv : by_n) groups.push_back(kv);
    for (int p : vols) groups.push_back({(int)probs_[p].g3->N, {p}});
    const size_t cb = 2 * esize();   // bytes of one complex element
    // vector layout: problems in class order so every class is contiguous
    int64_t off = 0, uoff = 0;
    int chunk = 0, slot = 0;
    // host tables stay alive until the uploads below have completed
    std::vector<std::vector<SliceRef>> keep;
    // classes
    for (auto &kv : groups) {
        auto k = std::make_unique<ToepClass>();
        const Geom3D *g3 = probs_[kv.second.front()].g3;
        k->N = kv.first;
        k->M = g3 ? (int)g3->M : (int)toep_conv_size(k->N);
        k->H = k->M / 2 + 1;
        LC_REQUIRE(fft_make_plan(k->M, k->plan), "batch: PSF size not 2^p 3^q");
        if (f64_) k->tw64 = c_.twiddles<double>(k->M);
        else k->tw = c_.twiddles<float>(k->M);
        if (g3) {
            // column blocks B * M <= the 3-D cap, of the block width the PSF
            // spectrum is stored in
            toep_configure(*k, toep3d_variant(k->M, f64_));
            LC_REQUIRE(k->B * k->M <= toep3d_col_cap(f64_), "batch: 3-D column block too large");
            LC_REQUIRE(k->B == (f64_ ? g3->kB64 : g3->kB),
                       "batch: 3-D PSF spectrum stored in another column block width");
            k->g3 = g3;
            k->n1 = (int)g3->n1;
            k->M1 = (int)g3->M1;
            LC_REQUIRE(fft_make_plan(k->M1, k->plan1), "batch: 3-D slice-axis size not 2^p 3^q");
            if (f64_) k->tw1_64 = c_.twiddles<double>(k->M1);
            else k->tw1 = c_.twiddles<float>(k->M1);
            LC_REQUIRE(k->B * k->M1 <= toep3d_slice_cap(f64_), "batch: 3-D slice axis too long");
            k->kb2 = std::max(1, toep3d_slice_cap(f64_) / (k->B * k->M1));
            k->W.alloc(cb * (size_t)k->nblk * k->M * k->n1 * k->B);
        } else {
            auto it = c_.toep_tuning.find(k->N);
            ToepVariant v = it != c_.toep_tuning.end() ? it->second : toep_default_variant(k->M);
            // a tuned FP32 column shape may not fit 16-byte elements
            if (v.col >= toep_num_col_shapes_for(f64_)) v = toep_default_variant(k->M);
            toep_configure(*k, toep_fit_variant(v, k->N, k->M, f64_));
        }
        std::vector<SliceRef> sl;
        for (int p : kv.second) {
            const int N = k->N, ns = probs_[p].ns;
            ProbDev &d = hprob_[p];
            d.khat = g3 ? nullptr
                        : (f64_ ? (const void *)probs_[p].g->khatT64.as()
                                : (const void *)probs_[p].g->khatT.as());
            d.N = N;
            d.ns = ns;
            d.off = off;
            d.n = (int64_t)ns * N * N;
            d.mu = 0.0f;
            d.boundary = 0;
            d.slot0 = slot;
            d.nslot = ns * k->p1_blocks;
            slot += d.nslot;
            d.chunk0 = chunk;
            d.nchunk = ceil_div(d.n, CHUNK);
            for (int i = 0; i < d.nchunk; ++i) {
                ChunkRef cr;
                cr.prob = p;
                cr.pad = 0;
                cr.begin = off + (int64_t)i * CHUNK;
                cr.end = std::min(off + d.n, cr.begin + CHUNK);
                hchunk_.push_back(cr);
            }
            chunk += d.nchunk;
            for (int s = 0; s < ns; ++s) {
                SliceRef r;
                r.prob = p;
                r.slice = s;
                r.xoff = off + (int64_t)s * N * N;
                r.uoff = uoff;
                uoff += toep_u_size(*k);
                sl.push_back(r);
            }
            // keep every problem 256-byte aligned (vector loads, coalescing)
            off += (d.n + 63) & ~(int64_t)63;
        }
        k->nslices = (int)sl.size();
        k->slot_base = hprob_[kv.second.front()].slot0;
        k->slices.alloc(sizeof(SliceRef) * sl.size());
        keep.push_back(std::move(sl));
        LC_HIP(hipMemcpyAsync(k->slices.as(), keep.back().data(),
                              sizeof(SliceRef) * keep.back().size(), hipMemcpyHostToDevice,
                              c_.stream()));
        classes_.push_back(std::move(k));
    }
    total_ = off;
    total_u_ = uoff;
    nchunk_ = chunk;
    nslot_ = slot;
    const size_t fb = esize() * (size_t)total_;
    X_.alloc(fb);
    R_.alloc(fb);
    P_.alloc(fb);
    AP_.alloc(fb);
    Y_.alloc(fb);
    U_.alloc(cb * (size_t)total_u_);
    dprob_.alloc(sizeof(ProbDev) * np);
    dchunk_.alloc(sizeof(ChunkRef) * hchunk_.size());
    dstate_.alloc(sizeof(CgState) * np);
    dsb_.alloc(sizeof(SbState) * np);
    partial_.alloc(sizeof(double) * std::max(1, nslot_));
    cpartial_.alloc(sizeof(double) * 4 * std::max(1, nchunk_));
    LC_HIP(hipMemcpyAsync(dprob_.as(), hprob_.data(), sizeof(ProbDev) * np, hipMemcpyHostToDevice,
                          c_.stream()));
    LC_HIP(hipMemcpyAsync(dchunk_.as(), hchunk_.data(), sizeof(ChunkRef) * hchunk_.size(),
                          hipMemcpyHostToDevice, c_.stream()));
    std::vector<CgState> z(np);
    for (auto &s : z) {
        std::memset(&s, 0, sizeof(s));
        s.active = 1;
    }
    LC_HIP(hipMemcpyAsync(dstate_.as(), z.data(), sizeof(CgState) * np, hipMemcpyHostToDevice,
                          c_.stream()));
    std::vector<SbState> zs(np);
    std::memset(zs.data(), 0, sizeof(SbState) * np);
    LC_HIP(hipMemcpyAsync(dsb_.as(), zs.data(), sizeof(SbState) * np, hipMemcpyHostToDevice,
                          c_.stream()));
    c_.sync();
}

// The Toeplitz passes of the size classes touch disjoint parts of the batch
// vectors, U and the slot partials, so the classes run on the context's main
// and auxiliary streams side by side (class i on stream i mod 4), joined back
// into the main stream before anything reads their results.
template <typename F> void Batch2D::for_classes(F &&launch) {
    const int nk = (int)classes_.size();
    const int ns = class_streams_ ? std::min(nk, 1 + Ctx::kAux) : 1;
    if (ns <= 1) {
        for (auto &k : classes_) launch(*k);
        return;
    }
    c_.fork(ns - 1);
    for (int i = 0; i < nk; ++i) {
        const int s = i % ns;
        Ctx::StreamScope sc(c_, s == 0 ? c_.main_stream() : c_.aux(s - 1));
        launch(*classes_[i]);
    }
    c_.join(ns - 1);
}

template <typename R> void Batch2D::apply_op_t(const R *src, R *dst, bool sb, bool dot) {
    CgState *st = dstate_.as<CgState>();
    cpx<R> *U = U_.as<cpx<R>>();
    for_classes([&](ToepClass &k) {
        toep_launch_fwd<R>(c_, k, st, src, nullptr, nullptr, U, false);
        toep_mid<R>(k, st, U);
        toep_launch_inv<R>(c_, k, st, U, dst, src, partial_.as<double>(), dot && !sb);
    });
    if (sb) {
        hipLaunchKernelGGL((k_sb_lapdot<R, 256>), dim3(nchunk_), dim3(256), 0, c_.stream(),
                           dchunk_.as<ChunkRef>(), dprob_.as<ProbDev>(), st, src, dst,
                           cpartial_.as<double>(), dot ? 1 : 0);
        LC_HIP(hipGetLastError());
    }
}

void Batch2D::apply_op(const void *src, void *dst, bool sb, bool dot) {
    LC_ETYPE([&](auto z) {
        using R = decltype(z);
        apply_op_t<R>(static_cast<const R *>(src), static_cast<R *>(dst), sb, dot);
    });
}

void Batch2D::toeplitz(const void *src, void *dst) {
    // all problems: arm every state active with no iterations
    std::vector<int> one(probs_.size(), 1);
    DBuf mi(sizeof(int) * one.size());
    LC_HIP(hipMemcpyAsync(mi.as(), one.data(), sizeof(int) * one.size(), hipMemcpyHostToDevice,
                          c_.stream()));
    hipLaunchKernelGGL(k_cg_arm, dim3(ceil_div(nprob(), 64)), dim3(64), 0, c_.stream(),
                       dstate_.as<CgState>(), nullptr, mi.as<int>(), 0.0, 0.0, nprob());
    LC_HIP(hipGetLastError());
    apply_op(src, dst, false, false);
    c_.sync();
}

float Batch2D::time_apply(int reps) {
    std::vector<int> one(probs_.size(), 1);
    DBuf mi(sizeof(int) * one.size());
    LC_HIP(hipMemcpyAsync(mi.as(), one.data(), sizeof(int) * one.size(), hipMemcpyHostToDevice,
                          c_.stream()));
    hipLaunchKernelGGL(k_cg_arm, dim3(ceil_div(nprob(), 64)), dim3(64), 0, c_.stream(),
                       dstate_.as<CgState>(), nullptr, mi.as<int>(), 0.0, 0.0, nprob());
    LC_ETYPE([&](auto z) {
        using R = decltype(z);
        hipLaunchKernelGGL(k_fill<R>, dim3(1024), dim3(256), 0, c_.stream(), X_.as<R>(), (R)1,
                           total_);
    });
    LC_HIP(hipGetLastError());
    apply_op(X_.as(), AP_.as(), false, false);   // warm-up
    hipEvent_t e0, e1;
    LC_HIP(hipEventCreate(&e0));
    LC_HIP(hipEventCreate(&e1));
    LC_HIP(hipEventRecord(e0, c_.stream()));
    for (int r = 0; r < reps; ++r) apply_op(X_.as(), AP_.as(), false, false);
    LC_HIP(hipEventRecord(e1, c_.stream()));
    LC_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    LC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c_.sync();
    return ms / (float)reps;
}

std::vector<Batch2D::PassTiming> Batch2D::profile_passes(int reps) {
    LC_REQUIRE(reps >= 1, "profile: reps must be >= 1");
    std::vector<int> one(probs_.size(), 1);
    DBuf mi(sizeof(int) * one.size());
    LC_HIP(hipMemcpyAsync(mi.as(), one.data(), sizeof(int) * one.size(), hipMemcpyHostToDevice,
                          c_.stream()));
    hipLaunchKernelGGL(k_cg_arm, dim3(ceil_div(nprob(), 64)), dim3(64), 0, c_.stream(),
                       dstate_.as<CgState>(), nullptr, mi.as<int>(), 0.0, 0.0, nprob());
    LC_HIP(hipGetLastError());
    CgState *st = dstate_.as<CgState>();
    const int nk = (int)classes_.size();
    std::vector<hipEvent_t> ev(4 * nk);
    for (auto &e : ev) LC_HIP(hipEventCreate(&e));
    std::vector<double> acc(3 * nk, 0.0);
    for (int r = 0; r <= reps; ++r) {   // r = 0 warms up
        for (int i = 0; i < nk; ++i) {
            const ToepClass &k = *classes_[i];
            LC_HIP(hipEventRecord(ev[4 * i], c_.stream()));
            LC_ETYPE([&](auto z) {
                using R = decltype(z);
                toep_launch_fwd<R>(c_, k, st, P_.as<R>(), R_.as<R>(), P_.as<R>(), U_.as<cpx<R>>(),
                                   true);
            });
            LC_HIP(hipEventRecord(ev[4 * i + 1], c_.stream()));
            LC_ETYPE([&](auto z) {
                using R = decltype(z);
                toep_mid<R>(k, st, U_.as<cpx<R>>());
            });
            LC_HIP(hipEventRecord(ev[4 * i + 2], c_.stream()));
            LC_ETYPE([&](auto z) {
                using R = decltype(z);
                toep_launch_inv<R>(c_, k, st, U_.as<cpx<R>>(), AP_.as<R>(), P_.as<R>(),
                                   partial_.as<double>(), true);
            });
            LC_HIP(hipEventRecord(ev[4 * i + 3], c_.stream()));
        }
        LC_HIP(hipEventSynchronize(ev[4 * nk - 1]));
        if (r == 0) continue;
        for (int i = 0; i < nk; ++i)
            for (int q = 0; q < 3; ++q) {
                float ms = 0.0f;
                LC_HIP(hipEventElapsedTime(&ms, ev[4 * i + q], ev[4 * i + q + 1]));
                acc[3 * i + q] += ms;
            }
    }
    // the vector update of a CG iteration (x += alpha p, r -= alpha Ap), with a
    // scratch buffer standing in for x so the solution is left untouched
    double upd_ms = 0.0;
    {
        DBuf xs(esize() * (size_t)total_);
        hipEvent_t u0, u1;
        LC_HIP(hipEventCreate(&u0));
        LC_HIP(hipEventCreate(&u1));
        for (int r = 0; r <= reps; ++r) {
            LC_HIP(hipEventRecord(u0, c_.stream()));
            LC_ETYPE([&](auto z) {
                using R = decltype(z);
                hipLaunchKernelGGL((k_cg_update<R, CG_UPD_NT>), dim3(nchunk_), dim3(CG_UPD_NT), 0, c_.stream(),
                                   dchunk_.as<ChunkRef>(), st, xs.as<R>(), R_.as<R>(), P_.as<R>(),
                                   AP_.as<R>(), cpartial_.as<double>());
            });
            LC_HIP(hipGetLastError());
            LC_HIP(hipEventRecord(u1, c_.stream()));
            LC_HIP(hipEventSynchronize(u1));
            float ms = 0.0f;
            LC_HIP(hipEventElapsedTime(&ms, u0, u1));
            if (r > 0) upd_ms += ms;
        }
        (void)hipEventDestroy(u0);
        (void)hipEventDestroy(u1);
        c_.sync();
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    std::vector<PassTiming> out;
    // element bytes e (4 or 8): real vectors e, complex 2e
    const double e = (double)esize();
    for (int i = 0; i < nk; ++i) {
        const ToepClass &k = *classes_[i];
        const double N = k.N, H = k.H, M = k.M, ns = k.nslices;
        double bytes[3] = {ns * (3 * e * N * N + 2 * e * N * H), ns * 4 * e * N * H + e * M * H,
                           ns * (2 * e * N * H + 2 * e * N * N)};
        if (k.g3)   // column FFTs to and from W, slice-axis pass over W, 3-D spectrum
            bytes[1] = ns * (4 * e * N * H + 4 * e * M * H) + 4 * e * M * H * ns + e * M * H * k.M1;
        for (int q = 0; q < 3; ++q)
            out.push_back(PassTiming{k.N, q, k.nslices, acc[3 * i + q] / reps, bytes[q]});
    }
    // pass 3: the update over every problem (reads p, Ap, x, r; writes x, r)
    int nsl = 0;
    for (auto &k : classes_) nsl += k->nslices;
    out.push_back(PassTiming{0, 3, nsl, upd_ms / reps, 6.0 * e * (double)total_});
    return out;
}

ToepVariant toep_autotune(Ctx &c, Geom2D &g, int nslices, double *best_ms) {
    if (!g.psf_ready) geom2d_build_psf(c, g);
    const int N = (int)g.N, M = (int)g.M;
    const auto saved = c.toep_tuning;
    auto time_variant = [&](const ToepVariant &v) -> double {
        c.toep_tuning[N] = v;
        Batch2D b(c, {ProbSpec{&g, nslices}});
        return (double)b.time_apply(3);
    };
    ToepVariant best = toep_default_variant(M);
    double tbest = 1e30;
    // stage 1: column pass shape and block width (row shape 0)
    for (int col = 0; col < toep_num_col_shapes(); ++col) {
        const int cap = toep_col_cap(col);
        if (M > cap) continue;
        const int bmax = std::max(1, std::min(cap / M, 16));
        int last = -1;
        for (int B : {bmax, bmax / 2, bmax / 4}) {
            if (B < 1 || B == last) continue;
            last = B;
            ToepVariant v;
            v.row = 0;
            v.col = col;
            v.B = B;
            try {
                const double t = time_variant(v);
                if (t < tbest) { tbest = t; best = v; }
            } catch (const Error &) {
                // shape not usable for this size
            }
        }
    }
    // stage 2: row pass shape with the chosen column shape
    for (int row = 1; row < toep_num_row_shapes(); ++row) {
        ToepVariant v = best;
        v.row = row;
        try {
            const double t = time_variant(v);
            if (t < tbest) { tbest = t; best = v; }
        } catch (const Error &) {
        }
    }
    c.toep_tuning = saved;
    c.toep_tuning[N] = best;
    if (best_ms) *best_ms = tbest;
    return best;
}

void Batch2D::cg_setup(bool x0_zero, bool sb, const std::vector<int> &max_iter, double tol,
                       double xtol) {
    (void)max_iter;
    (void)tol;
    (void)xtol;
    CgState *st = dstate_.as<CgState>();
    if (!x0_zero) apply_op(X_.as(), AP_.as(), sb, false);
    LC_ETYPE([&](auto z) {
        using R = decltype(z);
        hipLaunchKernelGGL((k_cg_init<R, 256>), dim3(nchunk_), dim3(256), 0, c_.stream(),
                           dchunk_.as<ChunkRef>(), st, Y_.as<R>(), AP_.as<R>(), R_.as<R>(),
                           P_.as<R>(), X_.as<R>(), cpartial_.as<double>(), x0_zero ? 1 : 0);
    });
    LC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cg_fin0, dim3(nprob()), dim3(FIN_NT), 0, c_.stream(), dprob_.as<ProbDev>(),
                       st, cpartial_.as<double>());
    LC_HIP(hipGetLastError());
}

template <typename R> void Batch2D::cg_iteration_t(bool host_sync, bool sb) {
    CgState *st = dstate_.as<CgState>();
    cpx<R> *U = U_.as<cpx<R>>();
    R *X = X_.as<R>(), *Rv = R_.as<R>(), *P = P_.as<R>(), *AP = AP_.as<R>();
    auto hs = [&]() {
        if (host_sync) {
            std::vector<CgState> h(nprob());
            LC_HIP(hipMemcpyAsync(h.data(), st, sizeof(CgState) * nprob(), hipMemcpyDeviceToHost,
                                  c_.stream()));
            c_.sync();
        }
    };
    for_classes([&](ToepClass &k) {
        toep_launch_fwd<R>(c_, k, st, P, Rv, P, U, true);
        toep_mid<R>(k, st, U);
        toep_launch_inv<R>(c_, k, st, U, AP, P, partial_.as<double>(), !sb);
    });
    if (sb) {
        hipLaunchKernelGGL((k_sb_lapdot<R, 256>), dim3(nchunk_), dim3(256), 0, c_.stream(),
                           dchunk_.as<ChunkRef>(), dprob_.as<ProbDev>(), st, P, AP,
                           cpartial_.as<double>(), 1);
        LC_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_cg_fin1, dim3(nprob()), dim3(FIN_NT), 0, c_.stream(), dprob_.as<ProbDev>(),
                       st, partial_.as<double>(), cpartial_.as<double>(), sb ? 1 : 0);
    LC_HIP(hipGetLastError());
    hs();
    hipLaunchKernelGGL((k_cg_update<R, CG_UPD_NT>), dim3(nchunk_), dim3(CG_UPD_NT), 0, c_.stream(),
                       dchunk_.as<ChunkRef>(), st, X, Rv, P, AP, cpartial_.as<double>());
    LC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cg_fin2, dim3(nprob()), dim3(FIN_NT), 0, c_.stream(), dprob_.as<ProbDev>(),
                       st, cpartial_.as<double>());
    LC_HIP(hipGetLastError());
    hs();
}

void Batch2D::cg_iteration(bool host_sync, bool sb) {
    LC_ETYPE([&](auto z) { cg_iteration_t<decltype(z)>(host_sync, sb); });
}

void Batch2D::cg(const std::vector<int> &max_iter, double tol, double xtol, bool graph,
                 bool host_sync, bool x0_zero) {
    LC_REQUIRE((int)max_iter.size() == nprob(), "cg: one max_iter per problem");
    const int iters = *std::max_element(max_iter.begin(), max_iter.end());
    // mu = 0: plain CG on A
    for (auto &d : hprob_) d.mu = 0.0f;
    LC_HIP(hipMemcpyAsync(dprob_.as(), hprob_.data(), sizeof(ProbDev) * nprob(),
                          hipMemcpyHostToDevice, c_.stream()));
    DBuf mi(sizeof(int) * max_iter.size());
    LC_HIP(hipMemcpyAsync(mi.as(), max_iter.data(), sizeof(int) * max_iter.size(),
                          hipMemcpyHostToDevice, c_.stream()));
    hipLaunchKernelGGL(k_cg_arm, dim3(ceil_div(nprob(), 64)), dim3(64), 0, c_.stream(),
                       dstate_.as<CgState>(), nullptr, mi.as<int>(), tol, xtol, nprob());
    LC_HIP(hipGetLastError());
    cg_setup(x0_zero, false, max_iter, tol, xtol);
    if (graph && !host_sync) {
        // CG_GRAPH_ITERS iterations per launch, then single ones; a problem
        // that stops (tolerance or its own max_iter) leaves the later
        // iterations' kernels with nothing to do
        if (!cg_graph1_) cg_graph1_ = std::make_unique<CapturedGraph>(c_, [&] { cg_iteration(false, false); });
        if (iters >= CG_GRAPH_ITERS && !cg_graphk_)
            cg_graphk_ = std::make_unique<CapturedGraph>(c_, [&] {
                for (int i = 0; i < CG_GRAPH_ITERS; ++i) cg_iteration(false, false);
            });
        int it = 0;
        for (; it + CG_GRAPH_ITERS <= iters; it += CG_GRAPH_ITERS) cg_graphk_->launch();
        for (; it < iters; ++it) cg_graph1_->launch();
    } else {
        for (int it = 0; it < iters; ++it) {
            cg_iteration(host_sync, false);
            if (host_sync) {
                // the plain path stops as soon as its (single) problem stops
                std::vector<CgState> h = cg_state();
                bool any = false;
                for (auto &s : h) any = any || s.active;
                if (!any) break;
            }
        }
    }
    c_.sync();
}

// one Split-Bregman outer iteration (src/bregman.cpp:58-88) for every active problem
template <typename R>
void Batch2D::sb_outer_t(bool hsync, double tol, double xtol, int inner, const int *dmi,
                         const std::vector<int> &mi) {
    const int np = nprob();
    const SbState *sbs = dsb_.as<SbState>();
    // rhs = y0 - mu div(d - b)
    hipLaunchKernelGGL((k_sb_rhs<R, 256>), dim3(nchunk_), dim3(256), 0, c_.stream(),
                       dchunk_.as<ChunkRef>(), dprob_.as<ProbDev>(), sbs, Y0_.as<R>(), Y_.as<R>(),
                       Dv_.as<R>(), Bv_.as<R>(), total_);
    LC_HIP(hipGetLastError());
    // x = cgsolver(A - mu lap, rhs, x, inner)
    hipLaunchKernelGGL(k_cg_arm, dim3(ceil_div(np, 64)), dim3(64), 0, c_.stream(),
                       dstate_.as<CgState>(), sbs, dmi, tol, xtol, np);
    cg_setup(false, true, mi, tol, xtol);
    for (int it = 0; it < inner; ++it) cg_iteration_t<R>(hsync, true);
    // shrink + Bregman update + convergence
    hipLaunchKernelGGL((k_sb_shrink<R, 256>), dim3(nchunk_), dim3(256), 0, c_.stream(),
                       dchunk_.as<ChunkRef>(), dprob_.as<ProbDev>(), sbs, X_.as<R>(), Dv_.as<R>(),
                       Bv_.as<R>(), XO_.as<R>(), cpartial_.as<double>(), total_,
                       sb_aniso_ ? 1 : 0);
    hipLaunchKernelGGL(k_sb_fin, dim3(np), dim3(FIN_NT), 0, c_.stream(), dprob_.as<ProbDev>(),
                       dsb_.as<SbState>(), cpartial_.as<double>());
    LC_HIP(hipGetLastError());
}

void Batch2D::split_bregman(const std::vector<float> &lambda, const std::vector<float> &mu,
                            int outer, int inner, double tol, double xtol, int boundary,
                            bool aniso, bool graph, bool host_sync, bool x0_zero) {
    const int np = nprob();
    LC_REQUIRE((int)lambda.size() == np && (int)mu.size() == np, "sb: one lambda/mu per problem");
    sb_boundary_ = boundary;
    sb_aniso_ = aniso;
    for (int p = 0; p < np; ++p) {
        hprob_[p].mu = mu[p];
        hprob_[p].boundary = boundary;
    }
    LC_HIP(hipMemcpyAsync(dprob_.as(), hprob_.data(), sizeof(ProbDev) * np, hipMemcpyHostToDevice,
                          c_.stream()));
    const size_t fb = esize() * (size_t)total_;
    Dv_.reserve(3 * fb);
    Bv_.reserve(3 * fb);
    XO_.reserve(fb);
    Y0_.reserve(fb);
    // d = b = 0, x_old = x0, y0 = A^T y
    LC_ETYPE([&](auto z) {
        using R = decltype(z);
        hipLaunchKernelGGL(k_fill<R>, dim3(1024), dim3(256), 0, c_.stream(), Dv_.as<R>(), (R)0,
                           3 * total_);
        hipLaunchKernelGGL(k_fill<R>, dim3(1024), dim3(256), 0, c_.stream(), Bv_.as<R>(), (R)0,
                           3 * total_);
        if (x0_zero)
            hipLaunchKernelGGL(k_fill<R>, dim3(1024), dim3(256), 0, c_.stream(), X_.as<R>(), (R)0,
                               total_);
    });
    LC_HIP(hipGetLastError());
    LC_HIP(hipMemcpyAsync(XO_.as(), X_.as(), fb, hipMemcpyDeviceToDevice, c_.stream()));
    LC_HIP(hipMemcpyAsync(Y0_.as(), Y_.as(), fb, hipMemcpyDeviceToDevice, c_.stream()));
    std::vector<SbState> hs(np);
    for (int p = 0; p < np; ++p) {
        std::memset(&hs[p], 0, sizeof(SbState));
        hs[p].lambda = lambda[p];
        hs[p].mu = mu[p];
        hs[p].xtol = xtol;
        hs[p].outer_max = outer;
        hs[p].active = outer > 0;
    }
    LC_HIP(hipMemcpyAsync(dsb_.as(), hs.data(), sizeof(SbState) * np, hipMemcpyHostToDevice,
                          c_.stream()));
    std::vector<int> mi(np, inner);
    DBuf dmi(sizeof(int) * np);
    LC_HIP(hipMemcpyAsync(dmi.as(), mi.data(), sizeof(int) * np, hipMemcpyHostToDevice,
                          c_.stream()));
    auto outer_iter = [&](bool hsync) {
        LC_ETYPE([&](auto z) {
            sb_outer_t<decltype(z)>(hsync, tol, xtol, inner, dmi.as<int>(), mi);
        });
    };
    if (graph && !host_sync) {
        CapturedGraph gr(c_, [&] { outer_iter(false); });
        for (int it = 0; it < outer; ++it) gr.launch();
    } else {
        for (int it = 0; it < outer; ++it) {
            outer_iter(host_sync);
            if (host_sync) {
                std::vector<SbState> h = sb_state();
                bool any = false;
                for (auto &s : h) any = any || s.active;
                if (!any) break;
            }
        }
    }
    // restore the CG rhs buffer to A^T y
    LC_HIP(hipMemcpyAsync(Y_.as(), Y0_.as(), fb, hipMemcpyDeviceToDevice, c_.stream()));
    c_.sync();
}

std::vector<CgState> Batch2D::cg_state() {
    std::vector<CgState> h(nprob());
    LC_HIP(hipMemcpyAsync(h.data(), dstate_.as(), sizeof(CgState) * nprob(), hipMemcpyDeviceToHost,
                          c_.stream()));
    c_.sync();
    return h;
}

std::vector<SbState> Batch2D::sb_state() {
    std::vector<SbState> h(nprob());
    LC_HIP(hipMemcpyAsync(h.data(), dsb_.as(), sizeof(SbState) * nprob(), hipMemcpyDeviceToHost,
                          c_.stream()));
    c_.sync();
    return h;
}

#undef LC_ETYPE

} // namespace lc

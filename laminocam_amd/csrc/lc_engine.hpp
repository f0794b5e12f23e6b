// 2-D geometry, Toeplitz normal operator and the batched CG / Split-Bregman
// engine.
#pragma once

#include <vector>

#include "lc_runtime.hpp"

namespace lc {

// ---------------------------------------------------------------------------
// NUFFT state of one 2-D polar grid: fine grid size, kernel parameters and
// the deconvolution factors (-1)^m / psi_hat(2 pi m / nf) for m in [-N/2, N/2)
// ---------------------------------------------------------------------------
struct Nufft2D {
    int nf = 0;           // fine grid edge (2N rounded up to 2^p 3^q)
    int w = 0;            // kernel width in cells
    double beta = 0.0;    // kernel shape
    DBuf corr;            // float[N]
    DBuf cs;              // double2[nang]: (cos theta, sin theta), FP64
    FftPlan plan_nf{};    // fine-grid FFT
    FftPlan plan_det{};   // detector FFT (valid if det_fast)
    bool det_fast = false;
    bool ready = false;
};

// ---------------------------------------------------------------------------
// Geometry of one 2-D polar grid (include/polar_grid2d.h:50-65) plus its
// Toeplitz PSF spectrum (include/toeplitz2d.h:81-99).
// ---------------------------------------------------------------------------
struct Geom2D {
    int64_t nang = 0, nrad = 0, N = 0, M = 0, H = 0;
    int64_t Mr = 0;    // the reference's PSF size next_fast_dim(2N - 1) (kernel_hat, psf_size);
                       // M = toep_conv_size(N) is the size the Toeplitz apply convolves at
    std::vector<double> theta;
    DBuf d_theta;      // double[nang]
    DBuf khatT;        // float[H][M]: real spectrum of the centred PSF / (M^2 N)
    DBuf khatT64;      // double[H][M]: the same spectrum for FP64 solves (built on demand)
    bool psf_ready = false, psf64_ready = false;
    Nufft2D nu;
};

// NUFFT projectors on device buffers (src/projection.cpp:146-182):
// forward2d: vol (ns x N x N) -> sino (ns x nang x nrad); backproj2d: reverse
void nufft2d_init(Ctx &c, Geom2D &g);
void forward2d_dev(Ctx &c, Geom2D &g, const float *vol, float *sino, int ns);
void backproj2d_dev(Ctx &c, Geom2D &g, const float *sino, float *vol, int ns);
// sysmat2d(x, grid) = Re(T1(T2 x)) / nrad (src/gradient.cpp:59-69)
void sysmat2d_nufft_dev(Ctx &c, Geom2D &g, const float *x, float *y, int ns);

// next_fast_dim of include/toeplitz2d.h:48-57
int64_t next_fast_dim(int64_t n);
// circular-convolution size of the Toeplitz apply for an n-point axis (see
// lc_toeplitz.hip); the reference's PSF size stays next_fast_dim(2 n - 1)
int64_t toep_conv_size(int64_t n);

// D_n(t) = sin(pi t) / sin(pi t / n) evaluated at t = k n + d as
// (-1)^{k(n-1)} sin(pi d) / sin(pi d / n), exact at the removable zeros
__device__ __forceinline__ double dirichlet_d(double t, int n, double inv_n) {
    const double kq = rint(t * inv_n);
    const double d = fma(-kq, (double)n, t);
    const long long k = (long long)kq;
    const double sg = ((k * (long long)(n - 1)) & 1LL) ? -1.0 : 1.0;
    if (d == 0.0) return sg * (double)n;
    return sg * sinpi(d) / sinpi(d * inv_n);
}

// ---------------------------------------------------------------------------
// Geometry of one 3-D laminography grid (include/polar_grid.h:52-94): per
// projection i the points are R_i (qX_k, qY_j, 0) with R_i =
// RotationTranspose(theta_i, gamma_i, beta_i) (include/rotation.h:28-56),
// qX_k = (k + 1/2) 2 pi / ncols - pi, qY_j = (j + 1/2) 2 pi / nrows - pi.
// Only the first two columns u_i, v_i of R_i are kept; points are generated
// on the fly.  Plus the Toeplitz PSF spectrum of include/toeplitz.h:61-129 for
// a {n1, N, N} volume.
// ---------------------------------------------------------------------------
// one detector axis of length n: Stockham plan, or Bluestein (plan of the
// padded length, chirp e^{i pi k^2/n}, chirp spectra for both signs / L)
struct DetLine {
    int n = 0;
    bool blue = false;
    FftPlan plan{};
    DBuf chirp, bhat;
};
void det_line_init(Ctx &c, DetLine &d, int n);

// NUFFT state of one 3-D grid: fine grid (z, y, x), per-axis corrections
struct Nufft3D {
    int nf[3] = {0, 0, 0};
    int coff[3] = {0, 0, 0};   // offsets of the per-axis factors in corr
    double beta = 0.0;
    DBuf corr;                 // float[n1 + n2 + n3]
    FftPlan plan[3]{};
    DetLine det_x, det_y;      // detector columns (ncols) and rows (nrows)
    bool ready = false;
};

struct Geom3D {
    int64_t nproj = 0, nrows = 0, ncols = 0;
    int64_t n1 = 0, n2 = 0, n3 = 0;   // recon dims
    int64_t N = 0;                    // n2 (== n3 for the Toeplitz operator)
    int64_t M1 = 0, M = 0, H = 0;     // convolution dims {M1, M, M}; H = M/2 + 1
    int64_t Mr = 0;                   // the reference's in-plane PSF size next_fast_dim(2N - 1)
    std::vector<double> rot;          // per projection: u (3) then v (3)
    DBuf d_rot;                       // double[nproj][6]
    // real spectrum of the centred PSF / (M1 M^2 N^2), in the column blocks of
    // the 3-D Toeplitz class: [ceil(H / kB)][M1][M][kB], value (k1, k2, k3) at
    // ((k3 / kB) M1 + k1) M kB + k2 kB + k3 % kB, zero for k3 >= H
    DBuf khat3;                       // float, block width kB
    DBuf khat3_64;                    // double, block width kB64 (FP64 solves, on demand)
    int kB = 0, kB64 = 0;
    bool psf_ready = false, psf64_ready = false;
    Nufft3D nu;
};

// 3-D projectors on device buffers (src/projection.cpp:42-131, src/gradient.cpp:32-41)
void nufft3d_init(Ctx &c, Geom3D &g);
void forward3d_dev(Ctx &c, Geom3D &g, const float *vol, float *proj);
void backproj3d_dev(Ctx &c, Geom3D &g, const float *proj, float *vol, const double *shifts);
void sysmat3d_nufft_dev(Ctx &c, Geom3D &g, const float *x, float *y);

void geom3d_init(Ctx &c, Geom3D &g, const double *theta, const double *gamma, const double *beta,
                 int64_t nproj, int64_t nrows, int64_t ncols, const int64_t *recon_dims);
// FP64 closed-form PSF P(m) = sum_i D_ncols(m.u_i) D_nrows(m.v_i), FP64 3-D FFT,
// stored as the FP32 real spectrum (f64: as the FP64 spectrum khat3_64)
void geom3d_build_psf(Ctx &c, Geom3D &g, bool f64 = false);
// reference-layout kernel_hat (complex {M1, M, M/2+1}, unscaled) to host, as
// float or double pairs
template <typename OT> void geom3d_kernel_hat(Ctx &c, Geom3D &g, OT *out_host);

void geom2d_init(Ctx &c, Geom2D &g, const double *theta, int64_t nang, int64_t nrad,
                 int64_t image_n);
// FP64 closed-form PSF + FP64 FFT, stored as FP32 real spectrum (f64: as the
// FP64 spectrum khatT64)
void geom2d_build_psf(Ctx &c, Geom2D &g, bool f64 = false);
// copy out the reference-layout kernel_hat (complex {M, M/2+1}, unscaled), as
// float or double pairs
template <typename OT> void geom2d_kernel_hat(Ctx &c, Geom2D &g, OT *out_host);

// ---------------------------------------------------------------------------
// device-side tables
// ---------------------------------------------------------------------------
struct SliceRef {
    int prob;       // problem index
    int slice;      // slice index within the problem
    int64_t xoff;   // float offset of the slice in the vector buffers
    int64_t uoff;   // complex offset of the slice's N x H intermediate
};

struct ProbDev {
    const void *khat;      // 2-D: spectrum [H][M] of the batch's element type
    int64_t off;    // float offset of the problem in the vector buffers
    int64_t n;      // ns * N * N
    int N, ns;
    int chunk0, nchunk;    // element chunks (update kernels)
    int slot0, nslot;      // Toeplitz partial-dot slots
    float mu;              // Split-Bregman coupling (0 for plain CG)
    int boundary;          // 0 interior-only (CPU reference), 1 zero padded
};

struct ChunkRef {
    int prob;
    int pad;
    int64_t begin, end;    // absolute float offsets
};

// per-problem CG scalars, FP64 (src/conjgrad.cpp:33-76)
struct CgState {
    double rho;      // <r, r>
    double pAp;
    double alpha;
    double beta;
    double ynorm;
    double res;
    double dx;
    double tol, xtol;
    int active;
    int iters;
    int stop;        // 0 running / max_iter, 1 tol, 2 xtol, 3 pAp guard
    int max_iter;
};

// Split-Bregman per-problem scalars (src/bregman.cpp:36-94)
struct SbState {
    double lambda, mu, xtol;
    double norm_diff;
    int active;
    int outer;
    int outer_max;
    int stop;
};

constexpr int CHUNK = 4096;   // elements per update-kernel block

// ---------------------------------------------------------------------------
// One Toeplitz size class: all slices that share N (hence M).
// ---------------------------------------------------------------------------
// shape of one class as the Toeplitz kernels see it (passed by value)
struct ToepGeom {
    int N, M, H;
    int B;      // columns per P2 block (U block width, a power of two)
    int lsh;    // log2(B)
    int nblk;   // ceil(H / B)
    int ppb;    // row pairs per P1/P3 block
};

struct ToepClass {
    int N = 0, M = 0, H = 0;
    // 3-D volume class (one problem): the middle pass becomes column FFT,
    // slice-axis convolution, inverse column FFT through W
    const Geom3D *g3 = nullptr;
    int n1 = 0, M1 = 0, kb2 = 1;      // slices, slice-axis FFT length, k2 rows per block
    FftPlan plan1{};
    const float2 *tw1 = nullptr;
    const double2 *tw1_64 = nullptr;
    DBuf W;                           // [nblk][M][n1][B] complex
    FftPlan plan{};
    const float2 *tw = nullptr;       // FFT_M twiddles, FP32 batches
    const double2 *tw64 = nullptr;    // FFT_M twiddles, FP64 batches
    int nslices = 0;
    DBuf slices;            // SliceRef[nslices]
    int ppb = 1, p1_blocks = 0;   // row pairs per block, blocks per slice
    int B = 1, nblk = 0, p2_blocks = 0;   // columns per block, blocks per slice
    int slot_base = 0;
    ToepGeom g{};
    ToepVariant var{};
};

template <typename R> inline const cpx<R> *toep_tw(const ToepClass &k);
template <> inline const float2 *toep_tw<float>(const ToepClass &k) { return k.tw; }
template <> inline const double2 *toep_tw<double>(const ToepClass &k) { return k.tw64; }

// block shapes for a class; U elements per slice (independent of the shapes)
void toep_configure(ToepClass &k, ToepVariant v);
ToepVariant toep_default_variant(int M);
// shapes of a 3-D volume class, and the column block width they give (the
// block width of the 3-D PSF spectrum's layout)
ToepVariant toep3d_variant(int M, bool f64);
int toep3d_block_width(int N, int M, bool f64);
// fine-grid sizes of the 2-D and 3-D NUFFTs (per axis), as their init sets them
int nufft2d_fine_n(int64_t N);
int nufft3d_fine_n(int64_t n);
size_t nufft3d_spread_bytes(int64_t npts, const int nf[3]);
// the 3-D spread's compiled bin shapes, and the autotuner that picks one per
// fine-grid shape (returns its index; installs it in c.nu3_tuning)
int nufft3d_num_shapes();
int nufft3d_autotune(Ctx &c, Geom3D &g, double *best_ms);
// the same for the 2-D spread, per fine-grid size, timed on nslices slices
int nufft2d_num_shapes();
int nufft2d_autotune(Ctx &c, Geom2D &g, int nslices, double *best_ms);
size_t nufft2d_spread_bytes(int64_t ns, int64_t npts, int nf);
// An odd image size N needs column blocks of at least two columns (the block
// copies move pairs of complex values).  A column shape too small for 2 M
// values is replaced by the first shape that holds them (block width then
// chosen by toep_configure).
ToepVariant toep_fit_variant(ToepVariant v, int N, int M, bool f64);
int64_t toep_u_size(const ToepClass &k);
int toep_num_row_shapes();
int toep_num_col_shapes();
int toep_col_cap(int col);
// largest column shape index usable with 8-byte (FP32) or 16-byte (FP64) complex
int toep_num_col_shapes_for(bool f64);
// time candidate shapes for image size g.N on `nslices` slices; installs the
// fastest in c.toep_tuning and returns it (best_ms: time of one apply)
ToepVariant toep_autotune(Ctx &c, Geom2D &g, int nslices, double *best_ms);

struct ProbSpec {
    Geom2D *g;             // 2-D: a stack of ns slices on one polar grid
    int ns;
    Geom3D *g3 = nullptr;  // 3-D: one {n1, N, N} volume (g == nullptr, ns == n1)
};

// middle Toeplitz pass of a 3-D volume class (three launches) and its
// block-size limits (B * M for the column FFTs, B * M1 for the slice axis)
template <typename R> void toep3d_mid(Ctx &c, const ToepClass &k, const CgState *st, cpx<R> *U);
int toep3d_col_cap(bool f64 = false);
int toep3d_slice_cap(bool f64 = false);

// the three Toeplitz passes of one class, element type R (float or double)
template <typename R>
void toep_launch_fwd(Ctx &c, const ToepClass &k, const CgState *st, const R *src, const R *rsrc,
                     R *pdst, cpx<R> *U, bool fuse);
template <typename R>
void toep_launch_cols(Ctx &c, const ToepClass &k, const ProbDev *pd, const CgState *st, cpx<R> *U);
template <typename R>
void toep_launch_inv(Ctx &c, const ToepClass &k, const CgState *st, const cpx<R> *U, R *dst,
                     const R *pvec, double *partial, bool dot);

// ---------------------------------------------------------------------------
// A batch of independent 2-D problems (each a stack of ns slices that share
// one geometry).  Vectors of all problems are concatenated; scalars are kept
// per problem so that a problem's iterates are the same alone and in a batch.
// f64 = true: vectors, spectra and FFTs are FP64 (the X() ... Y() pointers
// then address doubles; use the *v() accessors).
// ---------------------------------------------------------------------------
class Batch2D {
  public:
    // class_streams: run the size classes' Toeplitz passes on the context's
    // auxiliary streams side by side (off where several contexts share the
    // device's hardware queues, as the server's lanes do)
    Batch2D(Ctx &c, const std::vector<ProbSpec> &probs, bool f64 = false, bool class_streams = true);
    bool f64() const { return f64_; }
    size_t esize() const { return f64_ ? sizeof(double) : sizeof(float); }
    int nprob() const { return (int)probs_.size(); }
    int64_t total() const { return total_; }
    int64_t offset(int p) const { return hprob_[p].off; }
    int64_t size(int p) const { return hprob_[p].n; }
    const std::vector<ProbSpec> &specs() const { return probs_; }

    float *X() const { return X_.as<float>(); }
    float *R() const { return R_.as<float>(); }
    float *P() const { return P_.as<float>(); }
    float *AP() const { return AP_.as<float>(); }
    float *Y() const { return Y_.as<float>(); }
    // element-type-neutral views; elem(p, off) = address of element off of p
    void *Xv() const { return X_.as(); }
    void *Yv() const { return Y_.as(); }
    void *APv() const { return AP_.as(); }
    void *elem(void *p, int64_t off) const { return static_cast<char *>(p) + off * (int64_t)esize(); }

    // y = A x for every problem (Toeplitz only; src/dst are full vectors)
    void toeplitz(const void *src, void *dst);
    // average time (ms) of one Toeplitz apply over all problems (X -> AP)
    float time_apply(int reps);
    // HIP-event time of every launch of the three Toeplitz passes in the form a
    // CG iteration issues them (P1 fused with p = r + beta p, P3 with the
    // <p, Ap> partials), averaged over reps, then of the CG vector update over
    // the whole batch (pass 3, N = 0).  Overwrites P, AP, R and the solver
    // state, not X or Y.
    struct PassTiming {
        int N, pass, nslices;
        double ms;      // average duration of one launch
        double bytes;   // algorithmic HBM bytes of one launch
    };
    std::vector<PassTiming> profile_passes(int reps);

    // CG (src/conjgrad.cpp:33-76) on every problem, starting from X(),
    // right-hand side Y().  graph=true captures one iteration in a hipGraph
    // and replays it; host_sync=true reads the scalars back after each
    // kernel (the plain, one-solve-at-a-time mode).
    void cg(const std::vector<int> &max_iter, double tol, double xtol, bool graph,
            bool host_sync, bool x0_zero);
    // Split-Bregman TV (src/bregman.cpp:36-94) on every problem, starting from
    // X() (x0_zero: from x = 0, X() cleared first)
    void split_bregman(const std::vector<float> &lambda, const std::vector<float> &mu,
                       int outer, int inner, double tol, double xtol, int boundary,
                       bool aniso, bool graph, bool host_sync, bool x0_zero = false);

    std::vector<CgState> cg_state();
    std::vector<SbState> sb_state();

  private:
    void build();
    // one CG iteration (all kernels) for the current states
    void cg_iteration(bool host_sync, bool sb);
    template <typename R> void cg_iteration_t(bool host_sync, bool sb);
    void cg_setup(bool x0_zero, bool sb, const std::vector<int> &max_iter, double tol,
                  double xtol);
    void apply_op(const void *src, void *dst, bool sb, bool dot);
    template <typename R> void apply_op_t(const R *src, R *dst, bool sb, bool dot);
    template <typename F> void for_classes(F &&launch);
    // size classes on parallel streams (LC_CLASS_STREAMS=0 in the environment: one stream)
    bool class_streams_ = false;
    // middle pass of a class: 2-D column pass or the 3-D slice-coupled pass
    template <typename R> void toep_mid(const ToepClass &k, const CgState *st, cpx<R> *U);
    template <typename R> void sb_outer_t(bool hsync, double tol, double xtol, int inner,
                                          const int *dmi, const std::vector<int> &mi);

    Ctx &c_;
    bool f64_ = false;
    std::vector<ProbSpec> probs_;
    std::vector<ProbDev> hprob_;
    std::vector<ChunkRef> hchunk_;
    std::vector<std::unique_ptr<ToepClass>> classes_;
    int64_t total_ = 0, total_u_ = 0;
    int nchunk_ = 0, nslot_ = 0;
    DBuf X_, R_, P_, AP_, Y_, U_;
    DBuf dprob_, dchunk_, dstate_, dsb_, partial_, cpartial_;
    // Split-Bregman buffers (allocated on demand)
    DBuf Dv_, Bv_, XO_, Y0_;
    int sb_boundary_ = 0;
    bool sb_aniso_ = false;
    // executable graphs of one and of CG_GRAPH_ITERS CG iterations, captured
    // by the first graph-mode solve and replayed by later ones (their kernels
    // read every per-solve setting from device state); destroyed first
    std::unique_ptr<CapturedGraph> cg_graph1_, cg_graphk_;
};

// finite differences on one volume: op 0 grad (3 components out), 1 div
// (3 components in), 2 laplacian.  boundary 0 interior only, 1 zero padded.
void fd_launch(Ctx &c, int op, const float *in, float *out, int n1, int n2, int n3, int boundary);
void fd_launch(Ctx &c, int op, const double *in, double *out, int n1, int n2, int n3, int boundary);

// allocation-free CG (lc_noalloc.hip): noalloc_setup looks up the device
// address of the library's fixed scratch array on `device` (current) once;
// cg2d_noalloc runs lc_cg2d on device pointers with every device buffer
// carved from that array
void noalloc_setup(int device);
CgState cg2d_noalloc(Ctx &c, Geom2D &g, const float *rhs, float *x, int ns, int64_t max_iter,
                     double tol, double xtol, uint32_t flags);

// whole-solve drivers (lc_driver.hip), T = float or double
template <typename T>
void prepare_sinogram(Ctx &c, int32_t nds, const T *const *projs, const int64_t *nang, int64_t nrows,
                      int64_t ncols, T *sino, T *max_out, uint32_t flags);
template <typename T>
void mbir2d(Ctx &c, const T *sino, int64_t nslices, int64_t nang, int64_t ncols, const double *theta,
            const lc_recon_params *p, T *recon, lc_report *rep, uint32_t flags);
template <typename T>
void mbir3d(Ctx &c, int32_t nds, const T *const *projs, const int64_t *nang, int64_t nrows, int64_t ncols,
            const double *theta, const double *gamma, const double *beta, const double *shifts,
            const lc_recon_params *p, T *recon, int64_t *padded_dims, lc_report *rep, uint32_t flags);

} // namespace lc

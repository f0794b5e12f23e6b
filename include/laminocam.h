/*
 * laminocam.h - C ABI of the MI355X (gfx950) LaminoCam reconstruction hot path.
 *
 * Plain C: opaque handles, pointers and sizes, int status codes (0 = OK) and a
 * thread-local message from lc_last_error().  No C++ exception crosses this
 * boundary.  Arrays are row major float32; complex arrays interleave (re, im).
 * Entry points ending in _f64 take float64 arrays and run the whole operator
 * or solve in FP64 (vectors, PSF spectrum, FFTs and twiddles), standing in for
 * the reference's Array<double> instantiations.
 * Unless LC_DEVICE_PTRS is set, array arguments are host pointers and every
 * call completes before it returns.
 *
 * Each entry point names the reference interface it stands in for
 * (paths relative to the reference repository).
 */
#ifndef LAMINOCAM_H
#define LAMINOCAM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LC_ABI_VERSION 1

/* status codes */
#define LC_OK 0
#define LC_EINVAL 1
#define LC_EHIP 2
#define LC_ENOMEM 3
#define LC_EUNSUP 4
#define LC_EINTERNAL 5

/* flags */
#define LC_DEVICE_PTRS 0x1u  /* array arguments are device pointers */
#define LC_GRAPH 0x2u        /* replay solver iterations from a captured hipGraph */
#define LC_X0_ZERO 0x4u      /* start from x0 = 0: x is cleared, not read (skips A x0) */
#define LC_TV_ZERO_PAD 0x8u  /* TV differences zero padded (GPU reference) instead
                                of interior only (CPU reference) */
#define LC_TV_ANISO 0x10u    /* per-component soft threshold instead of isotropic */
#define LC_GEOM_LAZY_PSF 0x20u /* lc_geom2d_create: build the PSF on first use
                                 (geometries used only for projection) */
#define LC_FP64 0x40u        /* lc_batch_create_ex: FP64 vectors, spectra and FFTs */
#define LC_REQ_BORROW 0x80u  /* lc_request: sino / rhs / x0 are read in place when the request
                                runs instead of copied at submit; they must stay valid and
                                unchanged until the request finishes */
#define LC_HOST_SYNC 0x200u  /* lc_batch2d_cg / lc_batch2d_split_bregman: per-op launches that
                                read the solver scalars back after every iteration and stop
                                once every problem has stopped, the mode of the one-problem
                                entry points (without their allocations and copies) */
#define LC_REQ_TEST_FAIL 0x100u /* lc_request, a testing aid: the lane raises a host error when
                                   it builds a batch holding this request, before any device
                                   work, so that per-request failure isolation can be tested */

typedef struct lc_ctx_s *lc_ctx;
typedef struct lc_geom2d_s *lc_geom2d;
typedef struct lc_geom3d_s *lc_geom3d;
typedef struct lc_batch2d_s *lc_batch2d;

typedef struct {
    int64_t iters;  /* CG iterations run (last inner solve for Split-Bregman) */
    int64_t outer;  /* Split-Bregman outer iterations run */
    double res;     /* ||r|| / (||y|| + 1e-8) after the last CG iteration */
    double dx;      /* CG: ||alpha p||/(||x||+1e-8); SB: ||x-x_old||/(||x_old||+1e-8) */
    int32_t stop;   /* 0 iteration cap, 1 tol, 2 xtol, 3 |pAp| < 1e-10 */
    int32_t pad;
} lc_report;

int lc_abi_version(void);
const char *lc_last_error(void);

/* ---- context: one device, one stream ---- */
int lc_ctx_create(int device, lc_ctx *out);
/* drops the caller's handle; every geometry, batch and server created on the
 * context holds a share of it, and the context is freed (after waiting for its
 * stream) when the last of them is destroyed */
int lc_ctx_destroy(lc_ctx ctx);
int lc_ctx_sync(lc_ctx ctx);

/* ---- 2-D geometry + Toeplitz PSF ----
 * Replaces cpu::PolarGrid2D(theta, ncols) (include/polar_grid2d.h:50-65) and
 * cpu::PointSpreadFunction2D(grid, image_size) (include/toeplitz2d.h:81-99).
 * The PSF is computed in FP64 on the GPU from its closed form and stored as
 * an FP32 spectrum.  theta in radians. */
int lc_geom2d_create(lc_ctx ctx, const double *theta, int64_t nang, int64_t nrad,
                     int64_t image_n, uint32_t flags, lc_geom2d *out);
int lc_geom2d_destroy(lc_geom2d g);
/* info[0..4] = nang, nrad, N, psf_size M, M/2+1 */
int lc_geom2d_info(lc_geom2d g, int64_t *info);
/* PointSpreadFunction2D::kernel_hat() (include/toeplitz2d.h:63-66):
 * complex {M, M/2+1} written to out (host) */
int lc_psf2d_kernel_hat(lc_ctx ctx, lc_geom2d g, float *out);
/* PointSpreadFunction2D<double>::kernel_hat(): the same, as double pairs */
int lc_psf2d_kernel_hat_f64(lc_ctx ctx, lc_geom2d g, double *out);

/* ---- 2-D operators on a stack x of nslices N x N slices ----
 * sysmat2d(x, PointSpreadFunction2D) (src/gradient.cpp:76-79,
 * include/toeplitz2d.h:101-125) */
int lc_sysmat2d_psf(lc_ctx ctx, lc_geom2d g, const float *x, int64_t nslices, float *y,
                    uint32_t flags);
/* forward2d (src/projection.cpp:146-160, include/projection.h:56-58):
 * sino {nslices, nang, nrad} = Re(fftshift1(ifft1(ifftshift1(T2 vol))))/nrad
 * from vol {nslices, N, N}; T2 the type-2 NUFFT on the polar grid. */
int lc_forward2d(lc_ctx ctx, lc_geom2d g, const float *vol, int64_t nslices, float *sino,
                 uint32_t flags);
/* backproj2d (src/projection.cpp:168-182, include/projection.h:59-61):
 * vol = Re(T1(ifftshift1(fft1(fftshift1(sino)))))/nrad */
int lc_backproj2d(lc_ctx ctx, lc_geom2d g, const float *sino, int64_t nslices, float *vol,
                  uint32_t flags);
/* sysmat2d(x, PolarGrid2D) (src/gradient.cpp:59-69, include/projection.h:64-65):
 * Re(T1(T2 x))/nrad, the NUFFT form of the normal operator (no detector FFT) */
int lc_sysmat2d_nufft(lc_ctx ctx, lc_geom2d g, const float *x, int64_t nslices, float *y,
                      uint32_t flags);

/* ---- solvers on one stack (global alpha/beta over the stack, as the
 * reference's cgsolver on an Array{n1,N,N}) ----
 * opt::cgsolver(A = sysmat2d(., psf), y, x0, max_iter, tol, xtol)
 * (src/conjgrad.cpp:33-76).  x holds x0 on entry and the solution on exit. */
int lc_cg2d(lc_ctx ctx, lc_geom2d g, const float *rhs, float *x, int64_t nslices,
            int64_t max_iter, double tol, double xtol, uint32_t flags, lc_report *rep);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lc_cg2d with every device buffer of the solve (vectors,
 * Toeplitz intermediate, tables, partial sums) carved from a fixed array of
 * LC_NOALLOC_SCRATCH_MIB MiB that the library holds on each GPU until the
 * process ends; the call makes no hipMalloc or hipFree.  rhs and x are device
 * pointers (LC_DEVICE_PTRS is required); LC_X0_ZERO as lc_cg2d; LC_GRAPH is
 * refused.  The geometry must belong to ctx and have its PSF built.  A problem
 * whose buffers do not fit the array fails with LC_EINVAL before anything runs
 * (one 1024^2 slice, 8 slices of 512^2 or 32 of 256^2 fit).  Calls on one GPU
 * share the array: they run one at a time, each overwriting what the last
 * one left.  Same iterates and result as lc_cg2d. */
#define LC_NOALLOC_SCRATCH_MIB 64
int lc_cg2d_noalloc(lc_ctx ctx, lc_geom2d g, const float *rhs, float *x, int64_t nslices,
                    int64_t max_iter, double tol, double xtol, uint32_t flags, lc_report *rep);
/* opt::split_bregman (src/bregman.cpp:36-94) with A = sysmat2d(., psf) */
int lc_split_bregman2d(lc_ctx ctx, lc_geom2d g, const float *rhs, float *x, int64_t nslices,
                       double lambda, double mu, int64_t outer, int64_t inner, double tol,
                       double xtol, uint32_t flags, lc_report *rep);
/* FP64 forms: sysmat2d<double>(x, PointSpreadFunction2D<double>) (src/gradient.cpp:82-83),
 * cgsolver<double> (src/conjgrad.cpp:82-85), split_bregman<double> (src/bregman.cpp:101-104) */
int lc_sysmat2d_psf_f64(lc_ctx ctx, lc_geom2d g, const double *x, int64_t nslices, double *y,
                        uint32_t flags);
int lc_cg2d_f64(lc_ctx ctx, lc_geom2d g, const double *rhs, double *x, int64_t nslices,
                int64_t max_iter, double tol, double xtol, uint32_t flags, lc_report *rep);
int lc_split_bregman2d_f64(lc_ctx ctx, lc_geom2d g, const double *rhs, double *x,
                           int64_t nslices, double lambda, double mu, int64_t outer,
                           int64_t inner, double tol, double xtol, uint32_t flags,
                           lc_report *rep);

/* ---- 3-D laminography geometry + Toeplitz PSF ----
 * Replaces cpu::PolarGrid(theta, gamma, beta, nrows, ncols)
 * (include/polar_grid.h:52-94, rotation include/rotation.h:28-56) and
 * cpu::PointSpreadFunction(grid, recon_dims) (include/toeplitz.h:61-99) for a
 * volume recon_dims = {n1, n2, n3}.  The projectors take any dims; the
 * Toeplitz operator and the solvers need n2 == n3 (LC_EUNSUP otherwise).  Angles in radians, one (theta, gamma, beta) per projection.
 * The PSF is computed in FP64 on the GPU from its closed form
 * P(m) = sum_i D_ncols(m.u_i) D_nrows(m.v_i) and stored as an FP32 spectrum. */
int lc_geom3d_create(lc_ctx ctx, const double *theta, const double *gamma, const double *beta,
                     int64_t nproj, int64_t nrows, int64_t ncols, const int64_t *recon_dims,
                     uint32_t flags, lc_geom3d *out);
int lc_geom3d_destroy(lc_geom3d g);
/* info[0..8] = nproj, nrows, ncols, n1, n2, n3, PSF dims (3) */
int lc_geom3d_info(lc_geom3d g, int64_t *info);
/* PointSpreadFunction::kernel_hat() (include/toeplitz.h:55-57): complex
 * {d1, d2, d3/2+1} written to out (host) */
int lc_psf3d_kernel_hat(lc_ctx ctx, lc_geom3d g, float *out);
/* PointSpreadFunction<double>::kernel_hat(): the same, as double pairs */
int lc_psf3d_kernel_hat_f64(lc_ctx ctx, lc_geom3d g, double *out);
/* sysmat(x, PointSpreadFunction) (src/gradient.cpp:49-52, include/toeplitz.h:101-129)
 * on the {n1, n2, n3} volume x */
int lc_sysmat3d_psf(lc_ctx ctx, lc_geom3d g, const float *x, float *y, uint32_t flags);
/* forward(volume, grid) (src/projection.cpp:42-64, include/projection.h:38-39):
 * proj {nproj, nrows, ncols} = Re(ifftshift2(ifft2(fftshift2(w . T2 vol))))/(nrows ncols) */
int lc_forward3d(lc_ctx ctx, lc_geom3d g, const float *vol, float *proj, uint32_t flags);
/* backproj(proj, grid, recon_dims[, shifts]) (src/projection.cpp:71-131,
 * include/projection.h:41-47): Re(T1(w . phase . ifftshift2(fft2(fftshift2 proj))))/(nrows ncols).
 * shifts: NULL or nproj (dx, dy) pairs, phase = exp(-i (qx dx + qy dy)) (include/fft.h:53-74) */
int lc_backproj3d(lc_ctx ctx, lc_geom3d g, const float *proj, const double *shifts, float *vol,
                  uint32_t flags);
/* sysmat(x, PolarGrid) (src/gradient.cpp:32-41): Re(T1(T2 x))/(nrows ncols), no mask */
int lc_sysmat3d_nufft(lc_ctx ctx, lc_geom3d g, const float *x, float *y, uint32_t flags);
/* opt::cgsolver / opt::split_bregman with A = sysmat(., psf) on one volume
 * (src/conjgrad.cpp:33-76, src/bregman.cpp:36-94; the solve of src/mbir.cpp:122-157) */
int lc_cg3d(lc_ctx ctx, lc_geom3d g, const float *rhs, float *x, int64_t max_iter, double tol,
            double xtol, uint32_t flags, lc_report *rep);
int lc_split_bregman3d(lc_ctx ctx, lc_geom3d g, const float *rhs, float *x, double lambda,
                       double mu, int64_t outer, int64_t inner, double tol, double xtol,
                       uint32_t flags, lc_report *rep);
/* FP64 forms: sysmat<double>(x, PointSpreadFunction<double>) (src/gradient.cpp:55-56),
 * cgsolver<double>, split_bregman<double> */
int lc_sysmat3d_psf_f64(lc_ctx ctx, lc_geom3d g, const double *x, double *y, uint32_t flags);
int lc_cg3d_f64(lc_ctx ctx, lc_geom3d g, const double *rhs, double *x, int64_t max_iter,
                double tol, double xtol, uint32_t flags, lc_report *rep);
int lc_split_bregman3d_f64(lc_ctx ctx, lc_geom3d g, const double *rhs, double *x, double lambda,
                           double mu, int64_t outer, int64_t inner, double tol, double xtol,
                           uint32_t flags, lc_report *rep);

/* ---- finite differences on an n1 x n2 x n3 volume ----
 * opt::grad_u / divergence / laplacian (src/finitediff.cpp:26-90,
 * include/bregman.h:34-55).  g holds the 3 components back to back in the
 * order (slices, rows, cols). */
int lc_grad(lc_ctx ctx, const float *u, int64_t n1, int64_t n2, int64_t n3, float *g,
            uint32_t flags);
int lc_divergence(lc_ctx ctx, const float *g, int64_t n1, int64_t n2, int64_t n3, float *out,
                  uint32_t flags);
int lc_laplacian(lc_ctx ctx, const float *u, int64_t n1, int64_t n2, int64_t n3, float *out,
                 uint32_t flags);
/* grad_u<double> / divergence<double> (src/finitediff.cpp:69, 93) and the laplacian */
int lc_grad_f64(lc_ctx ctx, const double *u, int64_t n1, int64_t n2, int64_t n3, double *g,
                uint32_t flags);
int lc_divergence_f64(lc_ctx ctx, const double *g, int64_t n1, int64_t n2, int64_t n3,
                      double *out, uint32_t flags);
int lc_laplacian_f64(lc_ctx ctx, const double *u, int64_t n1, int64_t n2, int64_t n3, double *out,
                     uint32_t flags);

/* ---- batches of independent 2-D problems (per-problem alpha/beta) ----
 * No reference counterpart: gpu::MBIR2D (include/gpu/tomocam.h:60-75) solves
 * one stack at a time.  Problem i is a stack of ns[i] slices on geometry
 * geoms[i]; buffers stay resident on the device between calls. */
int lc_batch2d_create(lc_ctx ctx, int32_t nprob, const lc_geom2d *geoms, const int32_t *ns,
                      lc_batch2d *out);
/* mixed batch: problem i is a 2-D stack (g2[i], ns[i]) or a 3-D volume g3[i];
 * exactly one of g2[i], g3[i] is non-NULL (either array may be NULL) */
int lc_batch_create(lc_ctx ctx, int32_t nprob, const lc_geom2d *g2, const lc_geom3d *g3,
                    const int32_t *ns, lc_batch2d *out);
/* flags: LC_FP64 makes an FP64 batch, whose buffers are read and written with
 * the _f64 accessors below (the float accessors refuse it, and vice versa) */
int lc_batch_create_ex(lc_ctx ctx, int32_t nprob, const lc_geom2d *g2, const lc_geom3d *g3,
                       const int32_t *ns, uint32_t flags, lc_batch2d *out);
int lc_batch2d_destroy(lc_batch2d b);
/* which: 0 = x (solution / initial guess), 1 = y (right-hand side A^T y) */
int lc_batch2d_set(lc_batch2d b, int32_t prob, int32_t which, const float *src, uint32_t flags);
int lc_batch2d_get(lc_batch2d b, int32_t prob, int32_t which, float *dst, uint32_t flags);
/* device address of a problem's slot in buffer `which` */
int lc_batch2d_device_ptr(lc_batch2d b, int32_t prob, int32_t which, float **out);
int lc_batch2d_set_f64(lc_batch2d b, int32_t prob, int32_t which, const double *src,
                       uint32_t flags);
int lc_batch2d_get_f64(lc_batch2d b, int32_t prob, int32_t which, double *dst, uint32_t flags);
int lc_batch2d_device_ptr_f64(lc_batch2d b, int32_t prob, int32_t which, double **out);
int lc_batch2d_cg(lc_batch2d b, const int32_t *max_iter, double tol, double xtol,
                  uint32_t flags);
/* flags: LC_GRAPH, LC_TV_ZERO_PAD, LC_TV_ANISO, LC_X0_ZERO (start every problem from x = 0) */
int lc_batch2d_split_bregman(lc_batch2d b, const float *lambda, const float *mu, int64_t outer,
                             int64_t inner, double tol, double xtol, uint32_t flags);
int lc_batch2d_reports(lc_batch2d b, lc_report *reps);
/* y = A x over every problem of the batch (buffers 0 -> 1 unless stated) */
int lc_batch2d_sysmat(lc_batch2d b, int32_t src, int32_t dst);
/* Per-launch HIP-event timing of the three Toeplitz passes as one CG iteration
 * issues them, then of the CG vector update (no reference counterpart;
 * measurement only).  Writes one row of 5 doubles per (image size class,
 * pass): N, pass (0 rows+FFT, 1 columns, 2 inverse rows; 3 the update over the
 * whole batch, N = 0), slices, ms per launch, algorithmic HBM bytes per launch.
 * Leaves x and y untouched; clobbers the other solver buffers. */
int lc_batch2d_profile(lc_batch2d b, int32_t reps, int32_t max_rows, double *rows,
                       int32_t *nrows);

/* ---- reconstruction server (no reference counterpart) ----
 * Continuous batching with HBM-budget admission.  Requests queue on the host;
 * each of `lanes` device lanes (own stream, own thread) takes every queued
 * request that fits the free budget and shares the head request's solver
 * settings, runs them as one batch from captured graphs, and releases their
 * memory.  When the head request does not fit, smaller requests behind it
 * may start, for at most 8 batches; then the lanes wait until it fits.
 * Lanes overlap on the device.  Each problem keeps its own solver scalars:
 * its result is the same alone and in any mix.  A batch that fails is rerun
 * one request at a time, so only a request that fails on its own fails; a
 * request whose solution holds a NaN or an infinity fails too.  A failed
 * request reports its message through lc_server_result (LC_EINTERNAL). */
typedef struct lc_server_s *lc_server;
typedef struct {
    lc_geom2d g2;          /* a 2-D stack of nslices slices on g2, or */
    lc_geom3d g3;          /* one 3-D volume on g3 (exactly one is non-NULL) */
    int32_t nslices;
    int32_t solver;        /* 0 CG (src/conjgrad.cpp:33-76), 1 Split-Bregman (src/bregman.cpp:36-94) */
    int32_t max_iter;      /* CG iterations */
    int32_t outer, inner;  /* Split-Bregman iterations */
    double tol, xtol;
    double lambda, mu;     /* Split-Bregman */
    uint32_t flags;        /* LC_TV_ZERO_PAD, LC_TV_ANISO, LC_REQ_BORROW */
    const float *sino;     /* sinogram {nslices, nang, nrad} / projections {nproj, nrows, ncols},
                              back-projected on the device; or NULL and: */
    const float *rhs;      /* the right-hand side A^T y */
    const float *x0;       /* initial guess, NULL for zero */
    float *x_out;          /* NULL: the server keeps the solution until lc_server_result copies
                              it out; or a host buffer of the problem size that the lane writes
                              the solution into as the request finishes (valid until then) */
} lc_request;
int lc_server_create(lc_ctx ctx, uint64_t hbm_budget, int32_t lanes, int32_t max_batch,
                     lc_server *out);
/* copies the inputs (unless LC_REQ_BORROW); returns a ticket */
int lc_server_submit(lc_server s, const lc_request *r, uint64_t *ticket);
int lc_server_poll(lc_server s, uint64_t ticket, int32_t *done);
/* timeout_s < 0 waits without limit; *done = 0 on timeout */
int lc_server_wait(lc_server s, uint64_t ticket, double timeout_s, int32_t *done);
/* copies the solution into x (problem size) and releases the ticket; a request
 * submitted with x_out already holds its solution there, and x may be NULL (or
 * x_out: nothing is copied); times[3] = submit, start, done (seconds, steady clock) */
int lc_server_result(lc_server s, uint64_t ticket, float *x, lc_report *rep, double *times);
/* stats[10] = in-flight bytes, queued, running, finished, batches, budget, and
 * summed lane seconds building batches, loading inputs, solving, reading back */
int lc_server_stats(lc_server s, double *stats);
/* the most recent batches (up to max_rows, oldest first, the server keeps the
 * last 65536), 6 doubles each: lane, requests, host start and end (seconds,
 * steady clock), device start and end of the solve (ms since the server was
 * created, from HIP events on the lane's stream) */
int lc_server_timeline(lc_server s, int32_t max_rows, double *rows, int32_t *nrows);
/* finishes every queued request, then stops the lanes */
int lc_server_destroy(lc_server s);

/* ---- whole-solve drivers ----
 * ReconParams (include/recon_params.h:45-54): field for field, same types and
 * defaults (lc_recon_params_default).  The regularizer values follow
 * enum class Regularizer { SPLIT_BREGMAN, UNCONSTRAINED }. */
#define LC_REG_SPLIT_BREGMAN 0
#define LC_REG_UNCONSTRAINED 1
typedef struct {
    int32_t regularizer;    /* LC_REG_UNCONSTRAINED (CG) or LC_REG_SPLIT_BREGMAN */
    int32_t pad;
    int64_t recon_dims[3];  /* {n1, n2, n3}, n2 == n3 */
    int64_t max_iters;      /* 100: CG iterations, or Split-Bregman outer iterations */
    int64_t inner_iters;    /* 1: CG iterations per Split-Bregman outer iteration */
    float lambda;           /* 0.1 */
    float mu;               /* 10 */
    float tol;              /* 1e-5 */
    float xtol;             /* 1e-5 */
    float pad_factor;       /* 1.4142 (3-D driver only) */
    float pad2;
} lc_recon_params;
void lc_recon_params_default(lc_recon_params *p);

/* prepare_sinogram (include/mbir2d_prep.h:44-65): the ndatasets projection
 * stacks projs[d] {nang[d], nrows, ncols} are stacked, divided by their global
 * maximum (when it is > 0) and written as the sinogram {nrows, sum nang, ncols}
 * (one slice per detector row).  The maximum is returned in *max_out (may be
 * NULL).  The angles concatenate on the host.  LC_DEVICE_PTRS: projs[d], sino
 * and max_out are device pointers. */
int lc_prepare_sinogram(lc_ctx ctx, int32_t ndatasets, const float *const *projs,
                        const int64_t *nang, int64_t nrows, int64_t ncols, float *sino,
                        float *max_out, uint32_t flags);
int lc_prepare_sinogram_f64(lc_ctx ctx, int32_t ndatasets, const double *const *projs,
                            const int64_t *nang, int64_t nrows, int64_t ncols, double *sino,
                            double *max_out, uint32_t flags);

/* The 2-D slice-stack solve of cpu_mbir2d (src/mbir2d.cpp:66-100) and
 * gpu::MBIR2D (src/gpu/mbir2d.cu:36-85): PolarGrid2D(theta, ncols),
 * PointSpreadFunction2D(grid, recon_dims[2]), rhs = backproj2d of the first
 * recon_dims[0] slices of the prepared sinogram {nslices, nang, ncols}, then
 * CG or Split-Bregman (max_iters outer x inner_iters) from x0 = 0 on the whole
 * stack.  recon receives {recon_dims}.  recon_dims[0] <= nslices and
 * recon_dims[1] == recon_dims[2] (LC_EINVAL otherwise).  flags: LC_GRAPH,
 * LC_TV_ZERO_PAD, LC_TV_ANISO, LC_DEVICE_PTRS (sino and recon are device
 * pointers; theta stays on the host).  The _f64 form solves in FP64; its
 * back-projection runs in FP32, as every projector here. */
int lc_mbir2d(lc_ctx ctx, const float *sino, int64_t nslices, int64_t nang, int64_t ncols,
              const double *theta, const lc_recon_params *params, float *recon, lc_report *rep,
              uint32_t flags);
int lc_mbir2d_f64(lc_ctx ctx, const double *sino, int64_t nslices, int64_t nang, int64_t ncols,
                  const double *theta, const lc_recon_params *params, double *recon,
                  lc_report *rep, uint32_t flags);

/* MBIR (src/mbir.cpp:34-161), the 3-D laminography solve over ndatasets
 * datasets (Dataset_t, include/recon_params.h:32-39): projs[d] {nang[d],
 * nrows, ncols} with one gamma[d] and beta[d] per dataset; theta and shifts
 * ({dx, dy} per projection, may be NULL for none) concatenate over the
 * datasets.  Steps: stack, divide by the global maximum (when > 0), pad every
 * projection symmetrically with zeros to size_t(n * pad_factor), made odd by
 * -1 when even (skipped when pad_factor - 1 < 1e-6), grid and PSF on the
 * padded detector for the padded volume (each recon dim padded the same way),
 * rhs = backproj(..., shifts), CG or Split-Bregman from x0 = 0, symmetric crop
 * to recon_dims into recon.  padded_dims (may be NULL) receives {volume n1,
 * n2, n3, detector rows, cols}.  A pad factor that makes the padded volume
 * smaller than recon_dims is LC_EINVAL.  flags as lc_mbir2d; with
 * LC_DEVICE_PTRS, projs[d], shifts and recon are device pointers (theta,
 * gamma, beta stay on the host). */
int lc_mbir3d(lc_ctx ctx, int32_t ndatasets, const float *const *projs, const int64_t *nang,
              int64_t nrows, int64_t ncols, const double *theta, const double *gamma,
              const double *beta, const double *shifts, const lc_recon_params *params,
              float *recon, int64_t *padded_dims, lc_report *rep, uint32_t flags);
int lc_mbir3d_f64(lc_ctx ctx, int32_t ndatasets, const double *const *projs,
                  const int64_t *nang, int64_t nrows, int64_t ncols, const double *theta,
                  const double *gamma, const double *beta, const double *shifts,
                  const lc_recon_params *params, double *recon, int64_t *padded_dims,
                  lc_report *rep, uint32_t flags);

/* ---- memory model ----
 * lc_reference_memory_model: the reference's analytic device-memory model of
 * its single-device solvers, as a query (scripts/estimate_bregman_memory.py:18-62,
 * its fit rule src/gpu/mbir2d_multi.cu:459-470).  For a volume of N = n1 n2 n3
 * elements of elem_bytes bytes (4 or 8), out[5] =
 *   typical high-water mark   Split-Bregman 22 N, CG 11 N
 *   conservative mark         Split-Bregman 24 N, CG 11 N
 *   inputs                    2 N (right-hand side and initial guess)
 *   the fit rule's working set  Split-Bregman 27 N, CG 11 N
 *   N
 * all in bytes except N.  The script models Split-Bregman only; for CG the
 * reference states only the fit rule's 11 N, used for all three figures.
 * lc_volume_fits: the reference's fit rule (gpu::volume_fits_one_gpu,
 * include/gpu/tomocam.h:91-93): fit working set < 0.8 x free_bytes.
 * lc_volume_fits_one_gpu: the same with the free memory of the context's
 * device (hipMemGetInfo).
 * lc_request_bytes (no reference counterpart): the device bytes this
 * library's server charges against its HBM budget for a request (solver
 * vectors, Toeplitz intermediates, Split-Bregman fields, back-projection
 * scratch); what admission uses. */
int lc_reference_memory_model(const int64_t *dims, int32_t elem_bytes, int32_t regularizer,
                              double *out);
int lc_volume_fits(const lc_recon_params *params, int32_t elem_bytes, uint64_t free_bytes,
                   int32_t *fits);
int lc_volume_fits_one_gpu(lc_ctx ctx, const lc_recon_params *params, int32_t elem_bytes,
                           int32_t *fits);
int lc_request_bytes(const lc_request *r, uint64_t *bytes);

/* ---- Toeplitz kernel shapes per image size (no reference counterpart) ----
 * variant[0] = row-pass shape (0: 256 threads x 8 elements, 1: 512 x 8,
 * 2: 256 x 16), variant[1] = column-pass shape (0: 256 x 16, 1: 256 x 8,
 * 2: 512 x 16, FP32 only), variant[2] = columns per column-pass block (at
 * most 16, rounded down to a power of two; 0 = the widest that fits).
 * lc_autotune2d times the
 * candidate shapes on nslices slices of geometry g, installs the fastest for
 * g's image size in the context and returns it; batches created afterwards
 * use it.  lc_set_tuning2d installs a known choice. */
int lc_autotune2d(lc_ctx ctx, lc_geom2d g, int32_t nslices, int32_t *variant, double *ms);
int lc_set_tuning2d(lc_ctx ctx, int64_t image_n, const int32_t *variant);
int lc_get_tuning2d(lc_ctx ctx, int64_t image_n, int32_t *variant);

/* ---- 3-D NUFFT spread bin shapes per grid (no reference counterpart) ----
 * The type-1 spread of lc_backproj3d sorts the points into bins of footprint
 * corners; the library is compiled for several bin shapes (shape 0, 4 x 4 x 8
 * corner cells along z, y, x, is the default; 1 is 8 x 4 x 4, 2 is 8 x 8 x 8).
 * lc_autotune3d times a back-projection of g's grid with each shape,
 * installs the fastest in the context for g's fine-grid shape (the volume's
 * dims, oversampled) and returns it, with the time of one back-projection in
 * ms.  lc_set_tuning3d installs a known choice; lc_get_tuning3d reads the
 * current one.  Every choice gives the same result up to the order of the
 * spread's sums (it is deterministic for a given choice). */
int lc_autotune3d(lc_ctx ctx, lc_geom3d g, int32_t *shape, double *ms);
/* The 2-D spread of lc_backproj2d likewise (no reference counterpart): bins of
 * 16 x 16 (shape 0, the default) or 32 x 32 (shape 1) corner cells per fine-
 * grid size, timed on nslices slices of g.  Fine grids the sorted-bin spread
 * does not serve keep shape 0 and report 0 ms. */
int lc_autotune2d_nufft(lc_ctx ctx, lc_geom2d g, int32_t nslices, int32_t *shape, double *ms);
int lc_set_tuning2d_nufft(lc_ctx ctx, lc_geom2d g, int32_t shape);
int lc_get_tuning2d_nufft(lc_ctx ctx, lc_geom2d g, int32_t *shape);
int lc_set_tuning3d(lc_ctx ctx, lc_geom3d g, int32_t shape);
int lc_get_tuning3d(lc_ctx ctx, lc_geom3d g, int32_t *shape);

/* ---- file formats (host only; no context or device needed) ----
 * TIFF stacks replace include/tiff.h:31-117 (tomocam::tiff::read / write over
 * libtiff).  Samples are single 32-bit IEEE floats; pages are slices.
 * lc_tiff_info: page count and the (first) page's height and width; every
 * page must have that size.  Reads classic TIFF and BigTIFF, either byte
 * order, strips (not tiles), compression none, LZW, deflate or PackBits, no
 * predictor (LC_EUNSUP otherwise).
 * lc_tiff_read: the first `pages` pages, each cropped to its top-left rows x
 * cols (rows <= height, cols <= width), into out[pages][rows][cols].  The
 * reference's reader makes even heights and widths odd by dropping the last
 * row / column; its C++ and Python wrappers here pass those sizes.
 * lc_tiff_write: data[pages][rows][cols] as little-endian pages of one
 * uncompressed strip each, with the tags of the reference's writer
 * (MinIsBlack, top-left orientation, contiguous planes, IEEE float samples);
 * BigTIFF when the file reaches 4 GiB.
 * HDF5 replaces include/hdfread.h:34-96 (tomocam::h5::read_images /
 * read_angles over libhdf5), for numeric datasets: superblock versions 0-3,
 * symbol-table and compact-link groups, integer and IEEE float elements of
 * either byte order, contiguous / compact / chunked storage (v1 B-tree,
 * single-chunk, implicit and unpaged fixed-array indexes), filters deflate,
 * shuffle and Fletcher-32 (LC_EUNSUP otherwise).
 * lc_h5_shape: rank (<= 32) and dims of the dataset at the absolute path
 * `dataset` (for example "/images", "/coords/alpha").
 * lc_h5_read: all count elements, converted to float32 (f64 = 0, as the
 * reference's H5T_NATIVE_FLOAT reads) or float64 (f64 = 1), into out.
 * Errors: a missing file or object, a malformed file and a size mismatch are
 * LC_EINVAL. */
int lc_tiff_info(const char *path, int64_t *pages, int64_t *height, int64_t *width);
int lc_tiff_read(const char *path, float *out, int64_t pages, int64_t rows, int64_t cols);
int lc_tiff_write(const char *path, const float *data, int64_t pages, int64_t rows, int64_t cols);
int lc_h5_shape(const char *path, const char *dataset, int32_t *rank, int64_t *dims);
int lc_h5_read(const char *path, const char *dataset, void *out, int64_t count, int32_t f64);

#ifdef __cplusplus
}
#endif
#endif

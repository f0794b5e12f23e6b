/*
 * CPU oracle for the LaminoCam reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: the HIP product path
 * never links, loads or calls it.  Only tests/, the smoke check and the
 * cpu_baseline leg of bench.py use it.
 *
 * It restates, in plain C, the arithmetic the reference delegates to FINUFFT
 * and FFTW, plus the reference's own operator chains:
 *   polar grids          include/polar_grid2d.h:50-65, include/polar_grid.h:52-94,
 *                        include/rotation.h:28-56
 *   FFT + shifts         include/fft.h:53-184, include/fftutils.h:34-136
 *   NUFFT conventions    include/nufft.h:37-95 (type-1 iflag +1, type-2 iflag -1,
 *                        modes {ncols,nrows,nslices}, CMCL mode order)
 *   projectors           src/projection.cpp:42-182
 *   NUFFT normal op      src/gradient.cpp:32-69
 *   Toeplitz PSF         include/toeplitz2d.h:81-125, include/toeplitz.h:61-129
 *   CG                   src/conjgrad.cpp:33-76
 *   Split-Bregman        src/bregman.cpp:36-94
 *   finite differences   src/finitediff.cpp:26-90 (interior only) and
 *                        src/gpu/finitediff.cu:28-133 (zero padded)
 *
 * REAL is double for liboracle64 (the FP64 oracle) and float for liboracle32
 * (the FP32 CPU baseline, which mirrors the reference's float recon_tomo path).
 * Complex arrays are interleaved (re, im) pairs of REAL, row-major.
 */
#ifndef LAMINOCAM_ORACLE_H
#define LAMINOCAM_ORACLE_H

#include <stdint.h>

#ifndef REAL
#define REAL double
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- build info ---- */
int or_real_bytes(void);
int or_num_threads(void);
void or_set_num_threads(int n);

/* ---- FFT (unnormalised; sign -1 forward, +1 backward) ---- */
int64_t or_next_fast_dim(int64_t n);             /* include/toeplitz2d.h:48-57 */
int or_fft_lines(REAL *data, int64_t n, int64_t howmany, int64_t stride,
                 int64_t dist, int sign);         /* in place, complex */
int or_fft2(REAL *data, int64_t batch, int64_t n2, int64_t n3, int sign);
int or_fft3(REAL *data, int64_t n1, int64_t n2, int64_t n3, int sign);
void or_dft_naive(const REAL *in, REAL *out, int64_t n, int sign);

/* shifts with the reference's odd-length convention (include/fftutils.h)
 * axes: 1 -> last axis only, 2 -> last two axes.  inverse=0 fftshift,
 * inverse=1 ifftshift.  ncomp = 1 (real) or 2 (complex). */
void or_shift(const REAL *in, REAL *out, int64_t n1, int64_t n2, int64_t n3,
              int axes, int inverse, int ncomp);

/* ---- polar grids ---- */
void or_polar_grid2d(const REAL *theta, int64_t nang, int64_t nrad, REAL *x,
                     REAL *y);
void or_rotation_transpose(double theta, double gamma, double beta, double *R9);
void or_polar_grid3d(const REAL *theta, const REAL *gamma, const REAL *beta,
                     int64_t nproj, int64_t nrows, int64_t ncols, REAL *x,
                     REAL *y, REAL *z, REAL *w);

/* ---- non-uniform DFT (brute force) and NUFFT (ES kernel) ----
 * type 1: f[m] = sum_j c_j exp(+i m.w_j)   (c -> f)
 * type 2: c_j = sum_m f[m] exp(-i m.w_j)   (f -> c)
 * f is {nslices(3-D only), nrows, ncols} in CMCL order. */
void or_nudft2d(int type, const REAL *x, const REAL *y, int64_t npts, REAL *c,
                REAL *f, int64_t nrows, int64_t ncols);
void or_nudft3d(int type, const REAL *x, const REAL *y, const REAL *z,
                int64_t npts, REAL *c, REAL *f, int64_t nsl, int64_t nrows,
                int64_t ncols);
int or_nufft2d(int type, const REAL *x, const REAL *y, int64_t npts, REAL *c,
               REAL *f, int64_t nrows, int64_t ncols, double tol);
int or_nufft3d(int type, const REAL *x, const REAL *y, const REAL *z,
               int64_t npts, REAL *c, REAL *f, int64_t nsl, int64_t nrows,
               int64_t ncols, double tol);
int or_es_params(double tol, int *width, double *beta);

/* ---- 2-D projectors (src/projection.cpp:146-182, src/gradient.cpp:59-69) ----
 * tol <= 0 selects the brute-force NUDFT. */
int or_forward2d(const REAL *vol, int64_t ns, int64_t N, const REAL *theta,
                 int64_t nang, int64_t nrad, REAL *sino, double tol);
int or_backproj2d(const REAL *sino, int64_t ns, int64_t nang, int64_t nrad,
                  const REAL *theta, int64_t N, REAL *vol, double tol);
int or_sysmat2d_nufft(const REAL *x, int64_t ns, int64_t N, const REAL *theta,
                      int64_t nang, int64_t nrad, REAL *out, double tol);

/* ---- 3-D projectors (src/projection.cpp:42-131, src/gradient.cpp:32-41) ---- */
int or_forward3d(const REAL *vol, int64_t n1, int64_t n2, int64_t n3,
                 const REAL *theta, const REAL *gamma, const REAL *beta,
                 int64_t nproj, int64_t nrows, int64_t ncols, REAL *proj,
                 double tol);
int or_backproj3d(const REAL *proj, int64_t nproj, int64_t nrows, int64_t ncols,
                  const REAL *theta, const REAL *gamma, const REAL *beta,
                  const REAL *shifts, int64_t n1, int64_t n2, int64_t n3,
                  REAL *vol, double tol);
int or_sysmat3d_nufft(const REAL *x, int64_t n1, int64_t n2, int64_t n3,
                      const REAL *theta, const REAL *gamma, const REAL *beta,
                      int64_t nproj, int64_t nrows, int64_t ncols, REAL *out,
                      double tol);

/* ---- Toeplitz PSF ----
 * method 0: closed form P(m) = sum_i D_nrad(m . e_theta_i)  (exact)
 * method 1: type-1 NUFFT of ones (as the reference computes it) at tol. */
typedef struct or_psf or_psf;
or_psf *or_psf2d_create(const REAL *theta, int64_t nang, int64_t nrad,
                        int64_t image_n, int method, double tol);
or_psf *or_psf3d_create(const REAL *theta, const REAL *gamma, const REAL *beta,
                        int64_t nproj, int64_t nrows, int64_t ncols,
                        int64_t r1, int64_t r2, int64_t r3, int method,
                        double tol);
void or_psf_destroy(or_psf *p);
/* a PSF from given real values (CMCL order) on PSF dims d */
or_psf *or_psf_from_real(int dim, const int64_t *d, const REAL *P);
void or_psf_dims(const or_psf *p, int64_t *d3);   /* PSF dims (1,M,M) or 3-D */
int or_psf_real(const or_psf *p, REAL *out);      /* P in CMCL order */
int or_psf_kernel_hat(const or_psf *p, REAL *out);/* R2C half spectrum */
int or_psf_convolve(const or_psf *p, const REAL *x, int64_t n1, int64_t n2,
                    int64_t n3, REAL *out);
void or_psf_apply_op(void *psf_and_dims, const REAL *in, REAL *out);

/* ---- finite differences ----  boundary: 0 REF_CPU (interior only),
 * 1 ZERO_PAD (every point, zero outside). g holds 3 arrays back to back in
 * the order (slices, rows, cols). */
void or_grad(const REAL *u, int64_t n1, int64_t n2, int64_t n3, int boundary,
             REAL *g);
void or_div(const REAL *g, int64_t n1, int64_t n2, int64_t n3, int boundary,
            REAL *out);
void or_laplacian(const REAL *u, int64_t n1, int64_t n2, int64_t n3,
                  int boundary, REAL *out);

/* ---- solvers ---- */
typedef void (*or_op_fn)(void *ctx, const REAL *in, REAL *out);
typedef struct {
    int64_t iters;     /* CG iterations performed (last CG call for SB) */
    int64_t outer;     /* SB outer iterations performed */
    double res;        /* last residual ||r||/(||y||+1e-8) */
    double dx;         /* last ||alpha p||/(||x||+1e-8) or SB ||dx||/||x|| */
    int stop;          /* 0 max_iter, 1 tol, 2 xtol, 3 pAp guard */
} or_report;

int or_cgsolver(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                int64_t n, int64_t max_iter, double tol, double xtol,
                REAL *x_out, or_report *rep);
/* or_cgsolver that also keeps the iterates after iterations at[0..nat)
 * (ascending) in xs[j n ..]; an iterate past an early stop is the final x */
int or_cgsolver_trace(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                      int64_t n, int64_t max_iter, double tol, double xtol,
                      REAL *x_out, or_report *rep, const int64_t *at,
                      int64_t nat, REAL *xs);
int or_split_bregman(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                     int64_t n1, int64_t n2, int64_t n3, double lambda,
                     double mu, int64_t outer, int64_t inner, double tol,
                     double xtol, int boundary, int aniso, REAL *x_out,
                     or_report *rep);

/* the steps of one Split-Bregman outer iteration (src/bregman.cpp:58-84):
 * rhs = y - mu div(d - b); d = shrink(g + b, lambda/mu); b += g - d.
 * d, b, g hold 3 arrays of n values back to back. */
void or_sb_rhs(const REAL *y, const REAL *d, const REAL *b, int64_t n1,
               int64_t n2, int64_t n3, int boundary, double mu, REAL *rhs);
void or_sb_shrink(const REAL *g, const REAL *b, int64_t n, double lambda,
                  double mu, int aniso, REAL *d);
void or_sb_bregman_update(REAL *b, const REAL *g, const REAL *d, int64_t n);
/* or_split_bregman with caller-held d and b (3 n values each, in/out) */
int or_split_bregman_state(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                           int64_t n1, int64_t n2, int64_t n3, double lambda,
                           double mu, int64_t outer, int64_t inner, double tol,
                           double xtol, int boundary, int aniso, REAL *d,
                           REAL *b, REAL *x_out, or_report *rep);

/* convenience: Toeplitz operator context usable as or_op_fn */
typedef struct {
    const or_psf *psf;
    int64_t n1, n2, n3;
} or_psf_op_ctx;

#ifdef __cplusplus
}
#endif
#endif

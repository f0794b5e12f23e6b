/*
 * Operator chains of the reference restated for the oracle: polar grids,
 * projectors, NUFFT normal operators, Toeplitz PSFs, finite differences and
 * the CG / Split-Bregman solvers.  Test infrastructure only.
 */
#include "or_common.h"

#include <stdio.h>

/* ---------------- polar grids ---------------- */

/* include/polar_grid2d.h:50-65: x = cos(theta) r, y = sin(theta) r,
 * r_k = (k + 1/2) 2 pi / ncols - pi */
void or_polar_grid2d(const REAL *theta, int64_t nang, int64_t nrad, REAL *x,
                     REAL *y) {
    const double dr = 2.0 * M_PI / (double)nrad;
    for (int64_t i = 0; i < nang; ++i) {
        const double c = cos((double)theta[i]), s = sin((double)theta[i]);
        for (int64_t k = 0; k < nrad; ++k) {
            const double r = ((double)k + 0.5) * dr - M_PI;
            x[i * nrad + k] = (REAL)(c * r);
            y[i * nrad + k] = (REAL)(s * r);
        }
    }
}

/* include/rotation.h:28-56: transpose(Rx(theta) Ry(beta) Rz(gamma)), row major */
void or_rotation_transpose(double theta, double gamma, double beta, double *R) {
    const double ct = cos(theta), st = sin(theta);
    const double cb = cos(beta), sb = sin(beta);
    const double cg = cos(gamma), sg = sin(gamma);
    R[0] = cb * cg;
    R[1] = ct * sg + st * sb * cg;
    R[2] = st * sg - ct * sb * cg;
    R[3] = -cb * sg;
    R[4] = ct * cg - st * sb * sg;
    R[5] = st * cg + ct * sb * sg;
    R[6] = sb;
    R[7] = -st * cb;
    R[8] = ct * cb;
}

/* include/polar_grid.h:52-94: (x,y,z) = R^T(theta,gamma,beta) (qX, qY, 0),
 * w = 0 where any |coordinate| > pi */
void or_polar_grid3d(const REAL *theta, const REAL *gamma, const REAL *beta,
                     int64_t nproj, int64_t nrows, int64_t ncols, REAL *x,
                     REAL *y, REAL *z, REAL *w) {
    const double dX = 2.0 * M_PI / (double)ncols, dY = 2.0 * M_PI / (double)nrows;
    for (int64_t i = 0; i < nproj; ++i) {
        double R[9];
        or_rotation_transpose(theta[i], gamma[i], beta[i], R);
        for (int64_t j = 0; j < nrows; ++j)
            for (int64_t k = 0; k < ncols; ++k) {
                const double qX = ((double)k + 0.5) * dX - M_PI;
                const double qY = ((double)j + 0.5) * dY - M_PI;
                const double a = R[0] * qX + R[1] * qY;
                const double b = R[3] * qX + R[4] * qY;
                const double c = R[6] * qX + R[7] * qY;
                const int64_t q = (i * nrows + j) * ncols + k;
                x[q] = (REAL)a;
                y[q] = (REAL)b;
                z[q] = (REAL)c;
                w[q] = (fabs(a) > M_PI || fabs(b) > M_PI || fabs(c) > M_PI) ? 0 : 1;
            }
    }
}

/* ---------------- helpers ---------------- */

static REAL *to_cplx(const REAL *a, int64_t n) {
    REAL *c = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        c[2 * i] = a[i];
        c[2 * i + 1] = 0;
    }
    return c;
}

static int nufft2(int type, const REAL *x, const REAL *y, int64_t npts, REAL *c,
                  REAL *f, int64_t nr, int64_t nc, double tol) {
    if (tol <= 0) {
        or_nudft2d(type, x, y, npts, c, f, nr, nc);
        return 0;
    }
    return or_nufft2d(type, x, y, npts, c, f, nr, nc, tol);
}

static int nufft3(int type, const REAL *x, const REAL *y, const REAL *z,
                  int64_t npts, REAL *c, REAL *f, int64_t n1, int64_t n2,
                  int64_t n3, double tol) {
    if (tol <= 0) {
        or_nudft3d(type, x, y, z, npts, c, f, n1, n2, n3);
        return 0;
    }
    return or_nufft3d(type, x, y, z, npts, c, f, n1, n2, n3, tol);
}

/* ---------------- 2-D projectors ---------------- */

/* src/projection.cpp:146-160:
 * Re(fftshift1(ifft1(ifftshift1(nufft2d2(x)))))/nrad */
int or_forward2d(const REAL *vol, int64_t ns, int64_t N, const REAL *theta,
                 int64_t nang, int64_t nrad, REAL *sino, double tol) {
    const int64_t npts = nang * nrad;
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    or_polar_grid2d(theta, nang, nrad, x, y);
    REAL *C = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(ns * npts));
    REAL *T = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(ns * npts));
    for (int64_t s = 0; s < ns; ++s) {
        REAL *f = to_cplx(vol + s * N * N, N * N);
        nufft2(2, x, y, npts, C + 2 * s * npts, f, N, N, tol);
        free(f);
    }
    or_shift(C, T, 1, ns * nang, nrad, 1, 1, 2);
    or_fft_lines(T, nrad, ns * nang, 1, nrad, +1);
    or_shift(T, C, 1, ns * nang, nrad, 1, 0, 2);
    for (int64_t q = 0; q < ns * npts; ++q) sino[q] = (REAL)(C[2 * q] / (REAL)nrad);
    free(x);
    free(y);
    free(C);
    free(T);
    return 0;
}

/* src/projection.cpp:168-182:
 * Re(nufft2d1(ifftshift1(fft1(fftshift1(y)))))/nrad */
int or_backproj2d(const REAL *sino, int64_t ns, int64_t nang, int64_t nrad,
                  const REAL *theta, int64_t N, REAL *vol, double tol) {
    const int64_t npts = nang * nrad;
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    or_polar_grid2d(theta, nang, nrad, x, y);
    REAL *C = to_cplx(sino, ns * npts);
    REAL *T = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(ns * npts));
    or_shift(C, T, 1, ns * nang, nrad, 1, 0, 2);
    or_fft_lines(T, nrad, ns * nang, 1, nrad, -1);
    or_shift(T, C, 1, ns * nang, nrad, 1, 1, 2);
    REAL *F = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(N * N));
    for (int64_t s = 0; s < ns; ++s) {
        nufft2(1, x, y, npts, C + 2 * s * npts, F, N, N, tol);
        for (int64_t q = 0; q < N * N; ++q)
            vol[s * N * N + q] = (REAL)(F[2 * q] / (REAL)nrad);
    }
    free(x);
    free(y);
    free(C);
    free(T);
    free(F);
    return 0;
}

/* src/gradient.cpp:59-69: Re(T1(T2 x))/nrad, no FFT pair */
int or_sysmat2d_nufft(const REAL *xin, int64_t ns, int64_t N, const REAL *theta,
                      int64_t nang, int64_t nrad, REAL *out, double tol) {
    const int64_t npts = nang * nrad;
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    or_polar_grid2d(theta, nang, nrad, x, y);
    REAL *C = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)npts);
    for (int64_t s = 0; s < ns; ++s) {
        REAL *f = to_cplx(xin + s * N * N, N * N);
        nufft2(2, x, y, npts, C, f, N, N, tol);
        nufft2(1, x, y, npts, C, f, N, N, tol);
        for (int64_t q = 0; q < N * N; ++q) out[s * N * N + q] = (REAL)(f[2 * q] / (REAL)nrad);
        free(f);
    }
    free(x);
    free(y);
    free(C);
    return 0;
}

/* ---------------- 3-D projectors ---------------- */

/* src/projection.cpp:42-64:
 * Re(ifftshift2(ifft2(fftshift2(w . nufft3d2(x)))))/(n2 n3) */
int or_forward3d(const REAL *vol, int64_t n1, int64_t n2, int64_t n3,
                 const REAL *theta, const REAL *gamma, const REAL *beta,
                 int64_t nproj, int64_t nrows, int64_t ncols, REAL *proj,
                 double tol) {
    const int64_t npts = nproj * nrows * ncols;
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *z = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    or_polar_grid3d(theta, gamma, beta, nproj, nrows, ncols, x, y, z, w);
    REAL *f = to_cplx(vol, n1 * n2 * n3);
    REAL *C = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)npts);
    REAL *T = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)npts);
    nufft3(2, x, y, z, npts, C, f, n1, n2, n3, tol);
    for (int64_t q = 0; q < npts; ++q) {
        C[2 * q] *= w[q];
        C[2 * q + 1] *= w[q];
    }
    or_shift(C, T, nproj, nrows, ncols, 2, 0, 2);
    or_fft2(T, nproj, nrows, ncols, +1);
    or_shift(T, C, nproj, nrows, ncols, 2, 1, 2);
    const REAL sc = (REAL)(nrows * ncols);
    for (int64_t q = 0; q < npts; ++q) proj[q] = C[2 * q] / sc;
    free(x); free(y); free(z); free(w); free(f); free(C); free(T);
    return 0;
}

/* include/fft.h:53-74: C *= exp(-i (qx dx + qy dy)) per projection */
static void phase_shift2d(REAL *C, int64_t n1, int64_t n2, int64_t n3,
                          const REAL *shifts) {
    const double dqx = 2.0 * M_PI / (double)n3, dqy = 2.0 * M_PI / (double)n2;
    for (int64_t i = 0; i < n1; ++i) {
        const double dx = shifts[2 * i], dy = shifts[2 * i + 1];
        for (int64_t j = 0; j < n2; ++j)
            for (int64_t k = 0; k < n3; ++k) {
                const double qx = ((double)k + 0.5) * dqx - M_PI;
                const double qy = ((double)j + 0.5) * dqy - M_PI;
                const double ph = qx * dx + qy * dy;
                const double c = cos(ph), s = -sin(ph);
                const int64_t q = (i * n2 + j) * n3 + k;
                const double re = C[2 * q], im = C[2 * q + 1];
                C[2 * q] = (REAL)(re * c - im * s);
                C[2 * q + 1] = (REAL)(re * s + im * c);
            }
    }
}

/* src/projection.cpp:71-131:
 * Re(nufft3d1(w . [phase] ifftshift2(fft2(fftshift2(y)))))/(n2 n3) */
int or_backproj3d(const REAL *proj, int64_t nproj, int64_t nrows, int64_t ncols,
                  const REAL *theta, const REAL *gamma, const REAL *beta,
                  const REAL *shifts, int64_t n1, int64_t n2, int64_t n3,
                  REAL *vol, double tol) {
    const int64_t npts = nproj * nrows * ncols;
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *z = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    or_polar_grid3d(theta, gamma, beta, nproj, nrows, ncols, x, y, z, w);
    REAL *C = to_cplx(proj, npts);
    REAL *T = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)npts);
    or_shift(C, T, nproj, nrows, ncols, 2, 0, 2);
    or_fft2(T, nproj, nrows, ncols, -1);
    or_shift(T, C, nproj, nrows, ncols, 2, 1, 2);
    if (shifts) phase_shift2d(C, nproj, nrows, ncols, shifts);
    for (int64_t q = 0; q < npts; ++q) {
        C[2 * q] *= w[q];
        C[2 * q + 1] *= w[q];
    }
    REAL *F = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(n1 * n2 * n3));
    nufft3(1, x, y, z, npts, C, F, n1, n2, n3, tol);
    const REAL sc = (REAL)(nrows * ncols);
    for (int64_t q = 0; q < n1 * n2 * n3; ++q) vol[q] = F[2 * q] / sc;
    free(x); free(y); free(z); free(w); free(C); free(T); free(F);
    return 0;
}

/* src/gradient.cpp:32-41: Re(T1(T2 x))/(nrows ncols), no w mask */
int or_sysmat3d_nufft(const REAL *xin, int64_t n1, int64_t n2, int64_t n3,
                      const REAL *theta, const REAL *gamma, const REAL *beta,
                      int64_t nproj, int64_t nrows, int64_t ncols, REAL *out,
                      double tol) {
    const int64_t npts = nproj * nrows * ncols;
    REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *z = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
    or_polar_grid3d(theta, gamma, beta, nproj, nrows, ncols, x, y, z, w);
    REAL *f = to_cplx(xin, n1 * n2 * n3);
    REAL *C = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)npts);
    nufft3(2, x, y, z, npts, C, f, n1, n2, n3, tol);
    nufft3(1, x, y, z, npts, C, f, n1, n2, n3, tol);
    const REAL sc = (REAL)(nrows * ncols);
    for (int64_t q = 0; q < n1 * n2 * n3; ++q) out[q] = f[2 * q] / sc;
    free(x); free(y); free(z); free(w); free(f); free(C);
    return 0;
}

/* ---------------- Toeplitz PSF ---------------- */

/* D_n(t) = sum_k exp(i r_k t), r_k = (k+1/2) 2 pi/n - pi,
 *        = sin(pi t)/sin(pi t/n); at t = k n the limit n (-1)^{k(n-1)}.
 * Evaluated as (-1)^{k(n-1)} sin(pi d)/sin(pi d/n) with t = k n + d, |d| <= n/2,
 * so the ratio is never an indeterminate 0/0. */
static double dirichlet(double t, int64_t n) {
    const double kq = nearbyint(t / (double)n);
    const double d = t - kq * (double)n;
    const int64_t k = (int64_t)kq;
    const double sg = ((k * (n - 1)) & 1) ? -1.0 : 1.0;
    if (d == 0.0) return sg * (double)n;
    return sg * sin(M_PI * d) / sin(M_PI * d / (double)n);
}

struct or_psf {
    int dim;
    int64_t d[3];      /* PSF dims (slow -> fast); 2-D: {1, M, M} */
    REAL *P;           /* real PSF, CMCL order */
    REAL *K;           /* full complex FFT of P */
};

static void psf_finish(or_psf *p) {
    const int64_t n = p->d[0] * p->d[1] * p->d[2];
    p->K = to_cplx(p->P, n);
    if (p->dim == 3) or_fft3(p->K, p->d[0], p->d[1], p->d[2], -1);
    else or_fft2(p->K, 1, p->d[1], p->d[2], -1);
}

/* include/toeplitz2d.h:81-99 */
or_psf *or_psf2d_create(const REAL *theta, int64_t nang, int64_t nrad,
                        int64_t image_n, int method, double tol) {
    or_psf *p = (or_psf *)calloc(1, sizeof(or_psf));
    const int64_t M = or_next_fast_dim(2 * image_n - 1);
    p->dim = 2;
    p->d[0] = 1;
    p->d[1] = M;
    p->d[2] = M;
    p->P = (REAL *)malloc(sizeof(REAL) * (size_t)(M * M));
    if (method == 0) {
        double *ct = (double *)malloc(sizeof(double) * (size_t)nang);
        double *st = (double *)malloc(sizeof(double) * (size_t)nang);
        for (int64_t i = 0; i < nang; ++i) {
            ct[i] = cos((double)theta[i]);
            st[i] = sin((double)theta[i]);
        }
#pragma omp parallel for schedule(static)
        for (int64_t a = 0; a < M; ++a) {
            const double ky = (double)(a - M / 2);
            for (int64_t b = 0; b < M; ++b) {
                const double kx = (double)(b - M / 2);
                double s = 0;
                for (int64_t i = 0; i < nang; ++i) s += dirichlet(kx * ct[i] + ky * st[i], nrad);
                p->P[a * M + b] = (REAL)s;
            }
        }
        free(ct);
        free(st);
    } else {
        const int64_t npts = nang * nrad;
        REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
        REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
        or_polar_grid2d(theta, nang, nrad, x, y);
        REAL *ones = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)npts);
        for (int64_t q = 0; q < npts; ++q) {
            ones[2 * q] = 1;
            ones[2 * q + 1] = 0;
        }
        REAL *F = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(M * M));
        nufft2(1, x, y, npts, ones, F, M, M, tol);
        for (int64_t q = 0; q < M * M; ++q) p->P[q] = F[2 * q];
        free(x); free(y); free(ones); free(F);
    }
    psf_finish(p);
    return p;
}

/* include/toeplitz.h:61-91 (no w mask: the PSF uses every grid point) */
or_psf *or_psf3d_create(const REAL *theta, const REAL *gamma, const REAL *beta,
                        int64_t nproj, int64_t nrows, int64_t ncols,
                        int64_t r1, int64_t r2, int64_t r3, int method,
                        double tol) {
    or_psf *p = (or_psf *)calloc(1, sizeof(or_psf));
    p->dim = 3;
    p->d[0] = or_next_fast_dim(2 * r1 - 1);
    p->d[1] = or_next_fast_dim(2 * r2 - 1);
    p->d[2] = or_next_fast_dim(2 * r3 - 1);
    const int64_t D1 = p->d[0], D2 = p->d[1], D3 = p->d[2];
    p->P = (REAL *)malloc(sizeof(REAL) * (size_t)(D1 * D2 * D3));
    if (method == 0) {
        double *U = (double *)malloc(sizeof(double) * 6 * (size_t)nproj);
        for (int64_t i = 0; i < nproj; ++i) {
            double R[9];
            or_rotation_transpose(theta[i], gamma[i], beta[i], R);
            /* t1 = m . R[:,0] pairs with qX (ncols); t2 = m . R[:,1] with qY */
            U[6 * i + 0] = R[0]; U[6 * i + 1] = R[3]; U[6 * i + 2] = R[6];
            U[6 * i + 3] = R[1]; U[6 * i + 4] = R[4]; U[6 * i + 5] = R[7];
        }
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t s = 0; s < D1; ++s)
            for (int64_t a = 0; a < D2; ++a) {
                const double kz = (double)(s - D1 / 2), ky = (double)(a - D2 / 2);
                for (int64_t b = 0; b < D3; ++b) {
                    const double kx = (double)(b - D3 / 2);
                    double sum = 0;
                    for (int64_t i = 0; i < nproj; ++i) {
                        const double *u = U + 6 * i;
                        double t1 = kx * u[0] + ky * u[1] + kz * u[2];
                        double t2 = kx * u[3] + ky * u[4] + kz * u[5];
                        sum += dirichlet(t1, ncols) * dirichlet(t2, nrows);
                    }
                    p->P[(s * D2 + a) * D3 + b] = (REAL)sum;
                }
            }
        free(U);
    } else {
        const int64_t npts = nproj * nrows * ncols;
        REAL *x = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
        REAL *y = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
        REAL *z = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
        REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)npts);
        or_polar_grid3d(theta, gamma, beta, nproj, nrows, ncols, x, y, z, w);
        REAL *ones = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)npts);
        for (int64_t q = 0; q < npts; ++q) {
            ones[2 * q] = 1;
            ones[2 * q + 1] = 0;
        }
        REAL *F = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(D1 * D2 * D3));
        nufft3(1, x, y, z, npts, ones, F, D1, D2, D3, tol);
        for (int64_t q = 0; q < D1 * D2 * D3; ++q) p->P[q] = F[2 * q];
        free(x); free(y); free(z); free(w); free(ones); free(F);
    }
    psf_finish(p);
    return p;
}

/* a PSF from its real values P (CMCL order) on PSF dims d (d[0] = 1 with
 * dim 2); the operator then applies exactly as for a computed one */
or_psf *or_psf_from_real(int dim, const int64_t *d, const REAL *P) {
    or_psf *p = (or_psf *)calloc(1, sizeof(or_psf));
    p->dim = dim;
    p->d[0] = d[0];
    p->d[1] = d[1];
    p->d[2] = d[2];
    const int64_t n = d[0] * d[1] * d[2];
    p->P = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    memcpy(p->P, P, sizeof(REAL) * (size_t)n);
    psf_finish(p);
    return p;
}

void or_psf_destroy(or_psf *p) {
    if (!p) return;
    free(p->P);
    free(p->K);
    free(p);
}

void or_psf_dims(const or_psf *p, int64_t *d3) {
    d3[0] = p->d[0];
    d3[1] = p->d[1];
    d3[2] = p->d[2];
}

int or_psf_real(const or_psf *p, REAL *out) {
    memcpy(out, p->P, sizeof(REAL) * (size_t)(p->d[0] * p->d[1] * p->d[2]));
    return 0;
}

/* R2C layout {d0, d1, d2/2+1} of the full spectrum */
int or_psf_kernel_hat(const or_psf *p, REAL *out) {
    const int64_t H = p->d[2] / 2 + 1;
    for (int64_t s = 0; s < p->d[0] * p->d[1]; ++s)
        for (int64_t b = 0; b < H; ++b) {
            out[2 * (s * H + b)] = p->K[2 * (s * p->d[2] + b)];
            out[2 * (s * H + b) + 1] = p->K[2 * (s * p->d[2] + b) + 1];
        }
    return 0;
}

/* include/toeplitz2d.h:101-125 (2-D, per slice) and include/toeplitz.h:93-129
 * (3-D): pad RIGHT -> FFT -> x K -> IFFT -> crop at dims/2 -> / norm.
 * 2-D norm: M^2 * n3; 3-D norm: prod(dims) * n2 * n3 of the input. */
int or_psf_convolve(const or_psf *p, const REAL *x, int64_t n1, int64_t n2,
                    int64_t n3, REAL *out) {
    if (p->dim == 2) {
        const int64_t M = p->d[1];
        const double norm = (double)M * (double)M * (double)n3;
        REAL *buf = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)(M * M));
        for (int64_t s = 0; s < n1; ++s) {
            memset(buf, 0, sizeof(REAL) * 2 * (size_t)(M * M));
            for (int64_t a = 0; a < n2; ++a)
                for (int64_t b = 0; b < n3; ++b) buf[2 * (a * M + b)] = x[(s * n2 + a) * n3 + b];
            or_fft2(buf, 1, M, M, -1);
            for (int64_t q = 0; q < M * M; ++q) {
                const REAL re = buf[2 * q], im = buf[2 * q + 1];
                const REAL kr = p->K[2 * q], ki = p->K[2 * q + 1];
                buf[2 * q] = re * kr - im * ki;
                buf[2 * q + 1] = re * ki + im * kr;
            }
            or_fft2(buf, 1, M, M, +1);
            for (int64_t a = 0; a < n2; ++a)
                for (int64_t b = 0; b < n3; ++b)
                    out[(s * n2 + a) * n3 + b] =
                        (REAL)(buf[2 * ((a + M / 2) * M + b + M / 2)] / norm);
        }
        free(buf);
        return 0;
    }
    const int64_t D1 = p->d[0], D2 = p->d[1], D3 = p->d[2];
    const int64_t n = D1 * D2 * D3;
    const double norm = (double)n * (double)n2 * (double)n3;
    REAL *buf = (REAL *)calloc(2 * (size_t)n, sizeof(REAL));
    for (int64_t s = 0; s < n1; ++s)
        for (int64_t a = 0; a < n2; ++a)
            for (int64_t b = 0; b < n3; ++b)
                buf[2 * ((s * D2 + a) * D3 + b)] = x[(s * n2 + a) * n3 + b];
    or_fft3(buf, D1, D2, D3, -1);
    for (int64_t q = 0; q < n; ++q) {
        const REAL re = buf[2 * q], im = buf[2 * q + 1];
        const REAL kr = p->K[2 * q], ki = p->K[2 * q + 1];
        buf[2 * q] = re * kr - im * ki;
        buf[2 * q + 1] = re * ki + im * kr;
    }
    or_fft3(buf, D1, D2, D3, +1);
    for (int64_t s = 0; s < n1; ++s)
        for (int64_t a = 0; a < n2; ++a)
            for (int64_t b = 0; b < n3; ++b)
                out[(s * n2 + a) * n3 + b] = (REAL)(
                    buf[2 * (((s + D1 / 2) * D2 + a + D2 / 2) * D3 + b + D3 / 2)] / norm);
    free(buf);
    return 0;
}

void or_psf_apply_op(void *ctx, const REAL *in, REAL *out) {
    const or_psf_op_ctx *c = (const or_psf_op_ctx *)ctx;
    or_psf_convolve(c->psf, in, c->n1, c->n2, c->n3, out);
}

/* ---------------- finite differences ---------------- */

/* src/finitediff.cpp:26-63 (REF_CPU: interior points only, zero elsewhere);
 * ZERO_PAD: every point, values outside the volume read as zero
 * (src/gpu/finitediff.cu:28-133). */
void or_grad(const REAL *u, int64_t n1, int64_t n2, int64_t n3, int boundary,
             REAL *g) {
    const int64_t n = n1 * n2 * n3;
    memset(g, 0, sizeof(REAL) * 3 * (size_t)n);
#define U(i, j, k) u[((i) * n2 + (j)) * n3 + (k)]
    if (boundary == 0) {
        for (int64_t i = 1; i + 1 < n1; ++i)
            for (int64_t j = 1; j + 1 < n2; ++j)
                for (int64_t k = 1; k + 1 < n3; ++k) {
                    const int64_t q = (i * n2 + j) * n3 + k;
                    g[q] = (U(i + 1, j, k) - U(i - 1, j, k)) / (REAL)2.0;
                    g[n + q] = (U(i, j + 1, k) - U(i, j - 1, k)) / (REAL)2.0;
                    g[2 * n + q] = (U(i, j, k + 1) - U(i, j, k - 1)) / (REAL)2.0;
                }
    } else {
        for (int64_t i = 0; i < n1; ++i)
            for (int64_t j = 0; j < n2; ++j)
                for (int64_t k = 0; k < n3; ++k) {
                    const int64_t q = (i * n2 + j) * n3 + k;
                    REAL ip = (i + 1 < n1) ? U(i + 1, j, k) : 0, im = (i > 0) ? U(i - 1, j, k) : 0;
                    REAL jp = (j + 1 < n2) ? U(i, j + 1, k) : 0, jm = (j > 0) ? U(i, j - 1, k) : 0;
                    REAL kp = (k + 1 < n3) ? U(i, j, k + 1) : 0, km = (k > 0) ? U(i, j, k - 1) : 0;
                    g[q] = (REAL)0.5 * (ip - im);
                    g[n + q] = (REAL)0.5 * (jp - jm);
                    g[2 * n + q] = (REAL)0.5 * (kp - km);
                }
    }
#undef U
}

/* src/finitediff.cpp:66-87 */
void or_div(const REAL *g, int64_t n1, int64_t n2, int64_t n3, int boundary,
            REAL *out) {
    const int64_t n = n1 * n2 * n3;
    memset(out, 0, sizeof(REAL) * (size_t)n);
    const REAL *g0 = g, *g1 = g + n, *g2 = g + 2 * n;
#define I(i, j, k) (((i) * n2 + (j)) * n3 + (k))
    if (boundary == 0) {
        for (int64_t i = 1; i + 1 < n1; ++i)
            for (int64_t j = 1; j + 1 < n2; ++j)
                for (int64_t k = 1; k + 1 < n3; ++k)
                    out[I(i, j, k)] = (g0[I(i + 1, j, k)] - g0[I(i - 1, j, k)]) / (REAL)2.0 +
                                      (g1[I(i, j + 1, k)] - g1[I(i, j - 1, k)]) / (REAL)2.0 +
                                      (g2[I(i, j, k + 1)] - g2[I(i, j, k - 1)]) / (REAL)2.0;
    } else {
        for (int64_t i = 0; i < n1; ++i)
            for (int64_t j = 0; j < n2; ++j)
                for (int64_t k = 0; k < n3; ++k) {
                    REAL a = (REAL)0.5 * (((i + 1 < n1) ? g0[I(i + 1, j, k)] : 0) -
                                          ((i > 0) ? g0[I(i - 1, j, k)] : 0));
                    REAL b = (REAL)0.5 * (((j + 1 < n2) ? g1[I(i, j + 1, k)] : 0) -
                                          ((j > 0) ? g1[I(i, j - 1, k)] : 0));
                    REAL c = (REAL)0.5 * (((k + 1 < n3) ? g2[I(i, j, k + 1)] : 0) -
                                          ((k > 0) ? g2[I(i, j, k - 1)] : 0));
                    out[I(i, j, k)] = a + b + c;
                }
    }
#undef I
}

/* include/bregman.h:52-55 */
void or_laplacian(const REAL *u, int64_t n1, int64_t n2, int64_t n3,
                  int boundary, REAL *out) {
    const int64_t n = n1 * n2 * n3;
    REAL *g = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)n);
    or_grad(u, n1, n2, n3, boundary, g);
    or_div(g, n1, n2, n3, boundary, out);
    free(g);
}

/* ---------------- solvers ---------------- */

static double vdot(const REAL *a, const REAL *b, int64_t n) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return s;
}

/* src/conjgrad.cpp:33-76 (unpreconditioned; the CPU pAp guard at :53-56) */
/* or_cgsolver that also copies the iterate after iteration at[j] into
 * xs + j n for j < nat (at ascending; an iterate the loop does not reach,
 * because it stopped earlier, is the final x).  nat = 0: plain CG. */
int or_cgsolver_trace(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                      int64_t n, int64_t max_iter, double tol, double xtol,
                      REAL *x, or_report *rep, const int64_t *at, int64_t nat,
                      REAL *xs);

int or_cgsolver(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                int64_t n, int64_t max_iter, double tol, double xtol,
                REAL *x, or_report *rep) {
    return or_cgsolver_trace(A, ctx, y, x0, n, max_iter, tol, xtol, x, rep, NULL, 0, NULL);
}

int or_cgsolver_trace(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                      int64_t n, int64_t max_iter, double tol, double xtol,
                      REAL *x, or_report *rep, const int64_t *at, int64_t nat,
                      REAL *xs) {
    int64_t next = 0;   /* the next snapshot */
    REAL *r = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    REAL *p = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    REAL *Ap = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    memcpy(x, x0, sizeof(REAL) * (size_t)n);
    A(ctx, x, Ap);
    for (int64_t i = 0; i < n; ++i) r[i] = y[i] - Ap[i];
    const double ynorm = sqrt(vdot(y, y, n));
    memcpy(p, r, sizeof(REAL) * (size_t)n);
    double rs_old = vdot(r, r, n);
    or_report rp = {0, 0, 0, 0, 0};
    for (int64_t it = 0; it < max_iter; ++it) {
        A(ctx, p, Ap);
        const double pAp = vdot(p, Ap, n);
        if (fabs(pAp) < 1.0e-10) {
            rp.stop = 3;
            break;
        }
        const double alpha = rs_old / pAp;
        double step2 = 0;
        for (int64_t i = 0; i < n; ++i) {
            const REAL st = (REAL)(p[i] * alpha);
            step2 += (double)st * st;
            x[i] += st;
            r[i] -= (REAL)(Ap[i] * alpha);
        }
        const double rs_new = vdot(r, r, n);
        const double beta = rs_new / rs_old;
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + (REAL)(p[i] * beta);
        rs_old = rs_new;
        const double res = sqrt(rs_new) / (ynorm + 1e-8);
        const double dx = sqrt(step2) / (sqrt(vdot(x, x, n)) + 1e-8);
        rp.iters = it + 1;
        rp.res = res;
        rp.dx = dx;
        while (next < nat && at[next] <= it + 1) {
            if (at[next] == it + 1) memcpy(xs + next * n, x, sizeof(REAL) * (size_t)n);
            ++next;
        }
        if (res < tol) {
            rp.stop = 1;
            break;
        }
        if (dx < xtol) {
            rp.stop = 2;
            break;
        }
    }
    for (; next < nat; ++next)   /* snapshots past the stop: the final x */
        memcpy(xs + next * n, x, sizeof(REAL) * (size_t)n);
    if (rep) *rep = rp;
    free(r);
    free(p);
    free(Ap);
    return 0;
}

typedef struct {
    or_op_fn A;
    void *ctx;
    int64_t n1, n2, n3;
    int boundary;
    double mu;
    REAL *tmp;
} sb_ctx;

/* Ap(u) = A(u) - mu * laplacian(u)  (src/bregman.cpp:55) */
static void sb_op(void *c, const REAL *in, REAL *out) {
    sb_ctx *s = (sb_ctx *)c;
    const int64_t n = s->n1 * s->n2 * s->n3;
    s->A(s->ctx, in, out);
    or_laplacian(in, s->n1, s->n2, s->n3, s->boundary, s->tmp);
    for (int64_t i = 0; i < n; ++i) out[i] -= s->tmp[i] * (REAL)s->mu;
}

/* The three Split-Bregman steps of one outer iteration, each exposed so that
 * it can be checked on its own.
 *
 * rhs = yT - mu * div(d - b)                       src/bregman.cpp:58-61 */
void or_sb_rhs(const REAL *y, const REAL *d, const REAL *b, int64_t n1,
               int64_t n2, int64_t n3, int boundary, double mu, REAL *rhs) {
    const int64_t n = n1 * n2 * n3;
    REAL *g = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)n);
    for (int64_t i = 0; i < 3 * n; ++i) g[i] = d[i] - b[i];
    or_div(g, n1, n2, n3, boundary, rhs);
    for (int64_t i = 0; i < n; ++i) rhs[i] = y[i] - rhs[i] * (REAL)mu;
    free(g);
}

/* d from the gradient g of the new x and the Bregman variable b, n points of
 * 3 components each (src/bregman.cpp:66-80):
 *   isotropic:   sk = sqrt(sum_c (g_c + b_c)^2),
 *                d_c = max(0, sk - lambda/mu) * (g_c + b_c) / (sk + 1e-8)
 *   anisotropic: d_c = sign(v) max(0, |v| - lambda/mu), v = g_c + b_c
 *                (an option the reference does not have). */
void or_sb_shrink(const REAL *g, const REAL *b, int64_t n, double lambda,
                  double mu, int aniso, REAL *d) {
    const REAL thr = (REAL)(lambda / mu);
    for (int64_t i = 0; i < n; ++i) {
        if (!aniso) {
            double ss = 0;
            for (int c = 0; c < 3; ++c) {
                const REAL v = g[c * n + i] + b[c * n + i];
                ss += (double)v * v;
            }
            const REAL sk = (REAL)sqrt(ss);
            const REAL sh = (sk - thr > 0) ? sk - thr : 0;
            for (int c = 0; c < 3; ++c) {
                const REAL v = (g[c * n + i] + b[c * n + i]) / (sk + (REAL)1e-8);
                d[c * n + i] = sh * v;
            }
        } else {
            for (int c = 0; c < 3; ++c) {
                const REAL v = g[c * n + i] + b[c * n + i];
                const REAL av = v < 0 ? -v : v;
                const REAL sh = (av - thr > 0) ? av - thr : 0;
                d[c * n + i] = (v < 0) ? -sh : sh;
            }
        }
    }
}

/* Bregman update b += g - d over 3 n values (src/bregman.cpp:82-84) */
void or_sb_bregman_update(REAL *b, const REAL *g, const REAL *d, int64_t n) {
    for (int64_t i = 0; i < 3 * n; ++i) b[i] += g[i] - d[i];
}

/* src/bregman.cpp:36-94 with the splitting variables d and b (3 n values
 * each) supplied by the caller and updated in place; the reference starts
 * them at zero (src/bregman.cpp:44-48), as or_split_bregman does. */
int or_split_bregman_state(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                           int64_t n1, int64_t n2, int64_t n3, double lambda,
                           double mu, int64_t outer, int64_t inner, double tol,
                           double xtol, int boundary, int aniso, REAL *d,
                           REAL *b, REAL *x, or_report *rep) {
    const int64_t n = n1 * n2 * n3;
    REAL *xold = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    REAL *g = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)n);
    REAL *rhs = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    REAL *xn = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    sb_ctx s = {A, ctx, n1, n2, n3, boundary, mu,
                (REAL *)malloc(sizeof(REAL) * (size_t)n)};
    memcpy(x, x0, sizeof(REAL) * (size_t)n);
    memcpy(xold, x0, sizeof(REAL) * (size_t)n);
    or_report rp = {0, 0, 0, 0, 0}, cg;
    for (int64_t it = 0; it < outer; ++it) {
        or_sb_rhs(y, d, b, n1, n2, n3, boundary, mu, rhs);
        or_cgsolver(sb_op, &s, rhs, x, n, inner, tol, xtol, xn, &cg);
        memcpy(x, xn, sizeof(REAL) * (size_t)n);
        or_grad(x, n1, n2, n3, boundary, g);
        or_sb_shrink(g, b, n, lambda, mu, aniso, d);
        or_sb_bregman_update(b, g, d, n);
        double dn = 0, on = 0;
        for (int64_t i = 0; i < n; ++i) {
            const double df = (double)x[i] - (double)xold[i];
            dn += df * df;
            on += (double)xold[i] * xold[i];
        }
        const double nd = sqrt(dn) / (sqrt(on) + 1e-8);
        memcpy(xold, x, sizeof(REAL) * (size_t)n);
        rp.outer = it + 1;
        rp.iters = cg.iters;
        rp.res = cg.res;
        rp.dx = nd;
        if (nd < xtol) {
            rp.stop = 2;
            break;
        }
    }
    if (rep) *rep = rp;
    free(xold); free(g); free(rhs); free(xn); free(s.tmp);
    return 0;
}

/* src/bregman.cpp:36-94: d and b start at zero.  aniso=1 replaces the
 * isotropic shrink by the per-component soft threshold. */
int or_split_bregman(or_op_fn A, void *ctx, const REAL *y, const REAL *x0,
                     int64_t n1, int64_t n2, int64_t n3, double lambda,
                     double mu, int64_t outer, int64_t inner, double tol,
                     double xtol, int boundary, int aniso, REAL *x,
                     or_report *rep) {
    const int64_t n = n1 * n2 * n3;
    REAL *d = (REAL *)calloc(3 * (size_t)n, sizeof(REAL));
    REAL *b = (REAL *)calloc(3 * (size_t)n, sizeof(REAL));
    const int r = or_split_bregman_state(A, ctx, y, x0, n1, n2, n3, lambda, mu, outer,
                                         inner, tol, xtol, boundary, aniso, d, b, x, rep);
    free(d);
    free(b);
    return r;
}

/*
 * Non-uniform DFT (brute force) and exponential-of-semicircle (ES) NUFFT for
 * the oracle.  Conventions follow the reference's FINUFFT calls
 * (include/nufft.h:37-95): type 1 uses exp(+i m.w), type 2 exp(-i m.w); modes
 * are ordered {ncols, nrows, nslices} with CMCL order, i.e. array index
 * a <-> mode a - floor(n/2).
 *
 * The ES kernel and its Fourier transform are the published FINUFFT method
 * (Barnett, Magland, af Klinteberg, SIAM J. Sci. Comput. 41(5), 2019):
 * phi(z) = exp(beta (sqrt(1 - z^2) - 1)), width w cells, beta = 2.30 w for
 * upsampling 2; the correction 1/phihat is computed by Gauss-Legendre
 * quadrature (here after the substitution z = sin t, which removes the
 * square-root endpoint singularity).
 */
#include "or_common.h"

#include <stdio.h>

void or_es_choose(double tol, or_es *k) {
    if (tol < 1e-15) tol = 1e-15;
    int w = (int)ceil(log10(1.0 / tol)) + 1;
    if (w < 2) w = 2;
    if (w > 16) w = 16;
    k->w = w;
    k->beta = 2.30 * (double)w;
}

int or_es_params(double tol, int *width, double *beta) {
    or_es k;
    or_es_choose(tol, &k);
    *width = k.w;
    *beta = k.beta;
    return 0;
}

double or_es_phi(const or_es *k, double z) {
    if (z <= -1.0 || z >= 1.0) return 0.0;
    return exp(k->beta * (sqrt(1.0 - z * z) - 1.0));
}

/* Gauss-Legendre nodes on [-1, 1] by Newton iteration on P_n. */
#define GL_N 96
static double g_gl_x[GL_N], g_gl_w[GL_N];
static int g_gl_ready = 0;
static void gl_init(void) {
    if (g_gl_ready) return;
    const int n = GL_N;
    for (int i = 0; i < n; ++i) {
        double x = cos(M_PI * (i + 0.75) / (n + 0.5));
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
            if (fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = n * (x * p1 - p0) / (x * x - 1.0);
        g_gl_x[i] = x;
        g_gl_w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    g_gl_ready = 1;
}

/* Psi(xi) = int psi(u) e^{i xi u} du, psi(u) = phi(u / (w/2)), u in cells.
 *        = (w/2) int_{-pi/2}^{pi/2} e^{beta(cos t - 1)} cos(xi (w/2) sin t) cos t dt */
double or_es_phihat(const or_es *k, double xi) {
    gl_init();
    const double hw = 0.5 * k->w;
    double s = 0.0;
    for (int i = 0; i < GL_N; ++i) {
        double t = 0.5 * M_PI * g_gl_x[i];
        double ct = cos(t);
        s += g_gl_w[i] * exp(k->beta * (ct - 1.0)) * cos(xi * hw * sin(t)) * ct;
    }
    return hw * 0.5 * M_PI * s;
}

static int is_235(int64_t n) {
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

int64_t or_fine_size(int64_t nmodes, int w) {
    int64_t nf = 2 * nmodes;
    if (nf < 2 * w) nf = 2 * w;
    while (!(is_235(nf) && nf % 2 == 0)) ++nf;
    return nf;
}

/* ---------------- brute force NUDFT ---------------- */

void or_nudft2d(int type, const REAL *x, const REAL *y, int64_t npts, REAL *c,
                REAL *f, int64_t nrows, int64_t ncols) {
    const int64_t hr = nrows / 2, hc = ncols / 2;
    cx *cc = (cx *)c, *ff = (cx *)f;
    if (type == 1) {
#pragma omp parallel
        {
            double *ex = (double *)malloc(sizeof(double) * 2 * (size_t)ncols);
            double *ey = (double *)malloc(sizeof(double) * 2 * (size_t)nrows);
            double *acc = (double *)calloc(2 * (size_t)nrows * (size_t)ncols, sizeof(double));
#pragma omp for schedule(static)
            for (int64_t j = 0; j < npts; ++j) {
                for (int64_t b = 0; b < ncols; ++b) {
                    double a = (double)(b - hc) * (double)x[j];
                    ex[2 * b] = cos(a);
                    ex[2 * b + 1] = sin(a);
                }
                for (int64_t a = 0; a < nrows; ++a) {
                    double t = (double)(a - hr) * (double)y[j];
                    ey[2 * a] = cos(t);
                    ey[2 * a + 1] = sin(t);
                }
                const double cr = cc[j].re, ci = cc[j].im;
                for (int64_t a = 0; a < nrows; ++a) {
                    /* c * e^{i ky y} */
                    double pr = cr * ey[2 * a] - ci * ey[2 * a + 1];
                    double pi = cr * ey[2 * a + 1] + ci * ey[2 * a];
                    double *row = acc + 2 * a * ncols;
                    for (int64_t b = 0; b < ncols; ++b) {
                        row[2 * b] += pr * ex[2 * b] - pi * ex[2 * b + 1];
                        row[2 * b + 1] += pr * ex[2 * b + 1] + pi * ex[2 * b];
                    }
                }
            }
            /* merge per-thread partial sums in thread order (deterministic
             * for a fixed thread count) */
#ifdef _OPENMP
            int nt = omp_get_num_threads(), tid = omp_get_thread_num();
#else
            int nt = 1, tid = 0;
#endif
            for (int t = 0; t < nt; ++t) {
#pragma omp barrier
                if (t == tid) {
                    if (t == 0) {
                        for (int64_t q = 0; q < nrows * ncols; ++q)
                            ff[q] = cx_make((REAL)acc[2 * q], (REAL)acc[2 * q + 1]);
                    } else {
                        for (int64_t q = 0; q < nrows * ncols; ++q)
                            ff[q] = cx_make((REAL)(ff[q].re + acc[2 * q]),
                                            (REAL)(ff[q].im + acc[2 * q + 1]));
                    }
                }
            }
            free(ex);
            free(ey);
            free(acc);
        }
    } else {
#pragma omp parallel
        {
            double *ex = (double *)malloc(sizeof(double) * 2 * (size_t)ncols);
#pragma omp for schedule(static)
            for (int64_t j = 0; j < npts; ++j) {
                for (int64_t b = 0; b < ncols; ++b) {
                    double a = -(double)(b - hc) * (double)x[j];
                    ex[2 * b] = cos(a);
                    ex[2 * b + 1] = sin(a);
                }
                double sr = 0, si = 0;
                for (int64_t a = 0; a < nrows; ++a) {
                    double t = -(double)(a - hr) * (double)y[j];
                    double eyr = cos(t), eyi = sin(t);
                    const cx *row = ff + a * ncols;
                    double rr = 0, ri = 0;
                    for (int64_t b = 0; b < ncols; ++b) {
                        rr += row[b].re * ex[2 * b] - row[b].im * ex[2 * b + 1];
                        ri += row[b].re * ex[2 * b + 1] + row[b].im * ex[2 * b];
                    }
                    sr += rr * eyr - ri * eyi;
                    si += rr * eyi + ri * eyr;
                }
                cc[j] = cx_make((REAL)sr, (REAL)si);
            }
            free(ex);
        }
    }
}

void or_nudft3d(int type, const REAL *x, const REAL *y, const REAL *z,
                int64_t npts, REAL *c, REAL *f, int64_t nsl, int64_t nrows,
                int64_t ncols) {
    const int64_t hs = nsl / 2, hr = nrows / 2, hc = ncols / 2;
    cx *cc = (cx *)c, *ff = (cx *)f;
    const int64_t nm = nsl * nrows * ncols;
    if (type == 1) {
        /* serial over points, parallel over modes: deterministic */
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < nm; ++q) {
            int64_t s = q / (nrows * ncols), a = (q / ncols) % nrows, b = q % ncols;
            double kz = (double)(s - hs), ky = (double)(a - hr), kx = (double)(b - hc);
            double re = 0, im = 0;
            for (int64_t j = 0; j < npts; ++j) {
                double ph = kx * x[j] + ky * y[j] + kz * z[j];
                double cr = cos(ph), ci = sin(ph);
                re += cc[j].re * cr - cc[j].im * ci;
                im += cc[j].re * ci + cc[j].im * cr;
            }
            ff[q] = cx_make((REAL)re, (REAL)im);
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < npts; ++j) {
            double re = 0, im = 0;
            for (int64_t q = 0; q < nm; ++q) {
                int64_t s = q / (nrows * ncols), a = (q / ncols) % nrows, b = q % ncols;
                double ph = -((double)(b - hc) * x[j] + (double)(a - hr) * y[j] +
                              (double)(s - hs) * z[j]);
                double cr = cos(ph), ci = sin(ph);
                re += ff[q].re * cr - ff[q].im * ci;
                im += ff[q].re * ci + ff[q].im * cr;
            }
            cc[j] = cx_make((REAL)re, (REAL)im);
        }
    }
}

/* ---------------- ES-kernel NUFFT ---------------- */

/* kernel values for one coordinate: fills ker[0..w-1], returns first cell */
static inline int64_t es_weights(const or_es *k, double g, double *ker) {
    const double hw = 0.5 * k->w;
    int64_t l0 = (int64_t)ceil(g - hw);
    for (int i = 0; i < k->w; ++i) {
        double zz = ((double)(l0 + i) - g) / hw;
        ker[i] = or_es_phi(k, zz);
    }
    return l0;
}

static inline int64_t wrap(int64_t l, int64_t n) {
    l %= n;
    return l < 0 ? l + n : l;
}

/* grid coordinate in cells of a point in radians, periodic in [0, nf) */
static inline double to_cells(double x, int64_t nf) {
    double g = x * (double)nf / (2.0 * M_PI);
    double p = fmod(g, (double)nf);
    if (p < 0) p += (double)nf;
    return p;
}

/* general d-dimensional (2 or 3) NUFFT on fine grid nf[0..2] (slowest first) */
static int nufft_nd(int dim, int type, const REAL *const *pts, int64_t npts,
                    REAL *c, REAL *f, const int64_t *nm, double tol) {
    or_es k;
    or_es_choose(tol, &k);
    const int w = k.w;
    int64_t nf[3] = {1, 1, 1};
    int64_t mm[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
        mm[3 - dim + d] = nm[d];
        nf[3 - dim + d] = or_fine_size(nm[d], w);
    }
    /* pts are given as (x, y[, z]) : x <-> last axis, y <-> middle, z <-> first */
    const REAL *px = pts[0], *py = pts[1], *pz = (dim == 3) ? pts[2] : NULL;
    const int64_t nftot = nf[0] * nf[1] * nf[2];
    cx *fine = (cx *)calloc((size_t)nftot, sizeof(cx));
    if (!fine) return -1;
    cx *cc = (cx *)c, *ff = (cx *)f;

    /* correction factors per axis */
    double *corr[3];
    for (int d = 0; d < 3; ++d) {
        corr[d] = (double *)malloc(sizeof(double) * (size_t)mm[d]);
        for (int64_t a = 0; a < mm[d]; ++a) {
            if (mm[d] == 1 && nf[d] == 1) {
                corr[d][a] = 1.0;
                continue;
            }
            double kk = (double)(a - mm[d] / 2);
            corr[d][a] = 1.0 / or_es_phihat(&k, kk * 2.0 * M_PI / (double)nf[d]);
        }
    }

    if (type == 1) {
        /* spread (serial: fixed summation order) */
        double kx[16], ky[16], kz[16];
        for (int64_t j = 0; j < npts; ++j) {
            double gx = to_cells(px[j], nf[2]);
            double gy = to_cells(py[j], nf[1]);
            int64_t lx = es_weights(&k, gx, kx);
            int64_t ly = es_weights(&k, gy, ky);
            int64_t lz = 0;
            int wz = 1;
            kz[0] = 1.0;
            if (dim == 3) {
                double gz = to_cells(pz[j], nf[0]);
                lz = es_weights(&k, gz, kz);
                wz = w;
            }
            const double cr = cc[j].re, ci = cc[j].im;
            for (int iz = 0; iz < wz; ++iz) {
                int64_t zi = (dim == 3) ? wrap(lz + iz, nf[0]) : 0;
                for (int iy = 0; iy < w; ++iy) {
                    int64_t yi = wrap(ly + iy, nf[1]);
                    double wyz = ky[iy] * kz[iz];
                    cx *row = fine + (zi * nf[1] + yi) * nf[2];
                    for (int ix = 0; ix < w; ++ix) {
                        int64_t xi = wrap(lx + ix, nf[2]);
                        double wt = wyz * kx[ix];
                        row[xi].re += (REAL)(cr * wt);
                        row[xi].im += (REAL)(ci * wt);
                    }
                }
            }
        }
        if (dim == 3) or_fft3((REAL *)fine, nf[0], nf[1], nf[2], +1);
        else or_fft2((REAL *)fine, 1, nf[1], nf[2], +1);
        /* extract modes with correction */
#pragma omp parallel for collapse(2)
        for (int64_t s = 0; s < mm[0]; ++s)
            for (int64_t a = 0; a < mm[1]; ++a) {
                int64_t fs = wrap(s - mm[0] / 2, nf[0]);
                int64_t fa = wrap(a - mm[1] / 2, nf[1]);
                for (int64_t b = 0; b < mm[2]; ++b) {
                    int64_t fb = wrap(b - mm[2] / 2, nf[2]);
                    double cf = corr[0][s] * corr[1][a] * corr[2][b];
                    cx v = fine[(fs * nf[1] + fa) * nf[2] + fb];
                    ff[(s * mm[1] + a) * mm[2] + b] =
                        cx_make((REAL)(v.re * cf), (REAL)(v.im * cf));
                }
            }
    } else {
        /* place corrected modes on the fine grid */
#pragma omp parallel for collapse(2)
        for (int64_t s = 0; s < mm[0]; ++s)
            for (int64_t a = 0; a < mm[1]; ++a) {
                int64_t fs = wrap(s - mm[0] / 2, nf[0]);
                int64_t fa = wrap(a - mm[1] / 2, nf[1]);
                for (int64_t b = 0; b < mm[2]; ++b) {
                    int64_t fb = wrap(b - mm[2] / 2, nf[2]);
                    double cf = corr[0][s] * corr[1][a] * corr[2][b];
                    cx v = ff[(s * mm[1] + a) * mm[2] + b];
                    fine[(fs * nf[1] + fa) * nf[2] + fb] =
                        cx_make((REAL)(v.re * cf), (REAL)(v.im * cf));
                }
            }
        if (dim == 3) or_fft3((REAL *)fine, nf[0], nf[1], nf[2], -1);
        else or_fft2((REAL *)fine, 1, nf[1], nf[2], -1);
        /* interpolate */
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < npts; ++j) {
            double kx[16], ky[16], kz[16];
            double gx = to_cells(px[j], nf[2]);
            double gy = to_cells(py[j], nf[1]);
            int64_t lx = es_weights(&k, gx, kx);
            int64_t ly = es_weights(&k, gy, ky);
            int64_t lz = 0;
            int wz = 1;
            kz[0] = 1.0;
            if (dim == 3) {
                double gz = to_cells(pz[j], nf[0]);
                lz = es_weights(&k, gz, kz);
                wz = w;
            }
            double sr = 0, si = 0;
            for (int iz = 0; iz < wz; ++iz) {
                int64_t zi = (dim == 3) ? wrap(lz + iz, nf[0]) : 0;
                for (int iy = 0; iy < w; ++iy) {
                    int64_t yi = wrap(ly + iy, nf[1]);
                    double wyz = ky[iy] * kz[iz];
                    const cx *row = fine + (zi * nf[1] + yi) * nf[2];
                    double rr = 0, ri = 0;
                    for (int ix = 0; ix < w; ++ix) {
                        int64_t xi = wrap(lx + ix, nf[2]);
                        rr += row[xi].re * kx[ix];
                        ri += row[xi].im * kx[ix];
                    }
                    sr += rr * wyz;
                    si += ri * wyz;
                }
            }
            cc[j] = cx_make((REAL)sr, (REAL)si);
        }
    }
    for (int d = 0; d < 3; ++d) free(corr[d]);
    free(fine);
    return 0;
}

int or_nufft2d(int type, const REAL *x, const REAL *y, int64_t npts, REAL *c,
               REAL *f, int64_t nrows, int64_t ncols, double tol) {
    const REAL *pts[3] = {x, y, NULL};
    int64_t nm[2] = {nrows, ncols};
    return nufft_nd(2, type, pts, npts, c, f, nm, tol);
}

int or_nufft3d(int type, const REAL *x, const REAL *y, const REAL *z,
               int64_t npts, REAL *c, REAL *f, int64_t nsl, int64_t nrows,
               int64_t ncols, double tol) {
    const REAL *pts[3] = {x, y, z};
    int64_t nm[3] = {nsl, nrows, ncols};
    return nufft_nd(3, type, pts, npts, c, f, nm, tol);
}

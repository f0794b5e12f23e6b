/*
 * FFT restatement for the oracle: mixed-radix Stockham autosort for lengths
 * whose prime factors are <= 13, Bluestein (chirp-z through a power-of-two
 * plan) otherwise.  Unnormalised, sign -1 forward / +1 backward, matching
 * FFTW's FFTW_FORWARD / FFTW_BACKWARD as used by include/fft.h:76-184.
 */
#include "or_common.h"

#include <pthread.h>

/* next_fast_dim: smallest 2^p or 3^q 2^p >= n (include/toeplitz2d.h:48-57,
 * include/toeplitz.h:42-51). */
int64_t or_next_fast_dim(int64_t n) {
    int64_t best = 1;
    while (best < n) best <<= 1;
    for (int64_t q3 = 3; q3 < best; q3 *= 3) {
        int64_t x = q3;
        while (x < n) x <<= 1;
        if (x < best) best = x;
    }
    return best;
}

static int factorize(int64_t n, int *rad) {
    int nr = 0;
    while (n % 4 == 0 && n > 4) { rad[nr++] = 4; n /= 4; }
    const int primes[] = {2, 3, 5, 7, 11, 13};
    for (int i = 0; i < 6; ++i) {
        while (n % primes[i] == 0) { rad[nr++] = primes[i]; n /= primes[i]; }
    }
    if (n != 1) return -1;
    return nr;
}

static or_fftplan *plan_build(int64_t n) {
    or_fftplan *p = (or_fftplan *)calloc(1, sizeof(or_fftplan));
    p->n = n;
    p->tw = (cx *)malloc(sizeof(cx) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) p->tw[k] = cx_expi(-2.0 * M_PI * (double)k / (double)n);
    int nr = (n == 1) ? 0 : factorize(n, p->rad);
    if (nr >= 0) {
        p->nrad = nr;
        return p;
    }
    /* Bluestein */
    p->blue = 1;
    int64_t m = 1;
    while (m < 2 * n - 1) m <<= 1;
    p->m = m;
    p->sub = or_plan_get(m);
    p->chirp = (cx *)malloc(sizeof(cx) * (size_t)n);
    for (int64_t j = 0; j < n; ++j) {
        int64_t jj = (j * j) % (2 * n);
        p->chirp[j] = cx_expi(-M_PI * (double)jj / (double)n);
    }
    p->bfft = (cx *)calloc((size_t)m, sizeof(cx));
    p->bfft[0] = cx_conj(p->chirp[0]);
    for (int64_t j = 1; j < n; ++j) {
        p->bfft[j] = cx_conj(p->chirp[j]);
        p->bfft[m - j] = cx_conj(p->chirp[j]);
    }
    cx *work = (cx *)malloc(sizeof(cx) * or_plan_work(p->sub));
    or_plan_exec(p->sub, p->bfft, work, -1);
    free(work);
    return p;
}

/* ---- plan cache ---- */
#define MAX_PLANS 256
static or_fftplan *g_plans[MAX_PLANS];
static int g_nplans = 0;
static pthread_mutex_t g_plan_mu = PTHREAD_MUTEX_INITIALIZER;

or_fftplan *or_plan_get(int64_t n) {
    pthread_mutex_lock(&g_plan_mu);
    for (int i = 0; i < g_nplans; ++i) {
        if (g_plans[i]->n == n) {
            or_fftplan *p = g_plans[i];
            pthread_mutex_unlock(&g_plan_mu);
            return p;
        }
    }
    pthread_mutex_unlock(&g_plan_mu);
    /* build outside the lock (Bluestein recurses into or_plan_get) */
    or_fftplan *p = plan_build(n);
    pthread_mutex_lock(&g_plan_mu);
    for (int i = 0; i < g_nplans; ++i) {
        if (g_plans[i]->n == n) { /* lost the race: keep the first */
            or_fftplan *q = g_plans[i];
            pthread_mutex_unlock(&g_plan_mu);
            return q;
        }
    }
    if (g_nplans < MAX_PLANS) g_plans[g_nplans++] = p;
    pthread_mutex_unlock(&g_plan_mu);
    return p;
}

size_t or_plan_work(const or_fftplan *p) {
    if (p->blue) return (size_t)(4 * p->m + 2 * p->n);
    return (size_t)(2 * p->n);
}

/* One Stockham stage set.  Stage with radix R and Ns = product of previous
 * radices: for j < n/R, k = j mod Ns,
 *   v_r = in[j + r n/R] * w^{k r n/(Ns R)},  V = DFT_R(v),
 *   out[(j/Ns) Ns R + k + r Ns] = V_r. */
static void stockham(const or_fftplan *p, cx *a, cx *b, int sign) {
    const int64_t n = p->n;
    cx *in = a, *out = b;
    int64_t Ns = 1;
    cx v[16], V[16];
    for (int s = 0; s < p->nrad; ++s) {
        const int R = p->rad[s];
        const int64_t nR = n / R;
        const int64_t twstep = n / (Ns * R);
        for (int64_t j = 0; j < nR; ++j) {
            const int64_t k = j % Ns;
            for (int r = 0; r < R; ++r) {
                cx t = in[j + r * nR];
                if (k && r) {
                    cx w = p->tw[k * r * twstep];
                    if (sign > 0) w = cx_conj(w);
                    t = cx_mul(t, w);
                }
                v[r] = t;
            }
            if (R == 2) {
                V[0] = cx_add(v[0], v[1]);
                V[1] = cx_sub(v[0], v[1]);
            } else if (R == 4) {
                cx a0 = cx_add(v[0], v[2]), a1 = cx_sub(v[0], v[2]);
                cx b0 = cx_add(v[1], v[3]), b1 = cx_sub(v[1], v[3]);
                /* multiply b1 by sign*i */
                cx jb1 = (sign < 0) ? cx_make(b1.im, -b1.re) : cx_make(-b1.im, b1.re);
                V[0] = cx_add(a0, b0);
                V[2] = cx_sub(a0, b0);
                V[1] = cx_add(a1, jb1);
                V[3] = cx_sub(a1, jb1);
            } else {
                const int64_t st = n / R;
                for (int q = 0; q < R; ++q) {
                    cx acc = v[0];
                    for (int r = 1; r < R; ++r) {
                        cx w = p->tw[((int64_t)(q * r) % R) * st];
                        if (sign > 0) w = cx_conj(w);
                        acc = cx_add(acc, cx_mul(v[r], w));
                    }
                    V[q] = acc;
                }
            }
            const int64_t base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) out[base + r * Ns] = V[r];
        }
        cx *t = in;
        in = out;
        out = t;
        Ns *= R;
    }
    if (in != a) memcpy(a, in, sizeof(cx) * (size_t)n);
}

void or_plan_exec(const or_fftplan *p, cx *line, cx *work, int sign) {
    const int64_t n = p->n;
    if (n == 1) return;
    if (!p->blue) {
        stockham(p, line, work, sign);
        return;
    }
    /* Bluestein: X[k] = c_k sum_j (x_j c_j) conj(c_{k-j}), c_j = e^{s i pi j^2/n} */
    const int64_t m = p->m;
    cx *a = work;          /* m */
    cx *wk = work + m;     /* sub-plan work */
    for (int64_t j = 0; j < m; ++j) a[j] = cx_make(0, 0);
    for (int64_t j = 0; j < n; ++j) {
        cx c = p->chirp[j];
        if (sign > 0) c = cx_conj(c);
        a[j] = cx_mul(line[j], c);
    }
    or_plan_exec(p->sub, a, wk, -1);
    for (int64_t j = 0; j < m; ++j) {
        cx b = p->bfft[j];
        if (sign > 0) {
            /* FFT of conj(conj chirp) sequence = conj of FFT of reversed... use
             * symmetry: the + sequence is the conjugate of the - sequence,
             * whose FFT is conj(B[-k]). */
            b = cx_conj(p->bfft[(m - j) % m]);
        }
        a[j] = cx_mul(a[j], b);
    }
    or_plan_exec(p->sub, a, wk, +1);
    const REAL inv = (REAL)(1.0 / (double)m);
    for (int64_t k = 0; k < n; ++k) {
        cx c = p->chirp[k];
        if (sign > 0) c = cx_conj(c);
        line[k] = cx_scale(cx_mul(a[k], c), inv);
    }
}

int or_fft_lines(REAL *data, int64_t n, int64_t howmany, int64_t stride,
                 int64_t dist, int sign) {
    if (n <= 0 || howmany <= 0) return 0;
    or_fftplan *p = or_plan_get(n);
    cx *d = (cx *)data;
    const size_t wsz = or_plan_work(p) + (size_t)n;
#pragma omp parallel
    {
        cx *buf = (cx *)malloc(sizeof(cx) * wsz);
        cx *line = buf + or_plan_work(p);
#pragma omp for schedule(static)
        for (int64_t h = 0; h < howmany; ++h) {
            cx *base = d + h * dist;
            if (stride == 1) {
                or_plan_exec(p, base, buf, sign);
            } else {
                for (int64_t i = 0; i < n; ++i) line[i] = base[i * stride];
                or_plan_exec(p, line, buf, sign);
                for (int64_t i = 0; i < n; ++i) base[i * stride] = line[i];
            }
        }
        free(buf);
    }
    return 0;
}

int or_fft2(REAL *data, int64_t batch, int64_t n2, int64_t n3, int sign) {
    for (int64_t b = 0; b < batch; ++b) {
        REAL *s = data + 2 * b * n2 * n3;
        or_fft_lines(s, n3, n2, 1, n3, sign);   /* rows */
        or_fft_lines(s, n2, n3, n3, 1, sign);   /* columns */
    }
    return 0;
}

int or_fft3(REAL *data, int64_t n1, int64_t n2, int64_t n3, int sign) {
    or_fft_lines(data, n3, n1 * n2, 1, n3, sign);
    for (int64_t i = 0; i < n1; ++i)
        or_fft_lines(data + 2 * i * n2 * n3, n2, n3, n3, 1, sign);
    or_fft_lines(data, n1, n2 * n3, n2 * n3, 1, sign);
    return 0;
}

void or_dft_naive(const REAL *in, REAL *out, int64_t n, int sign) {
    const cx *x = (const cx *)in;
    cx *y = (cx *)out;
    for (int64_t k = 0; k < n; ++k) {
        double re = 0, im = 0;
        for (int64_t j = 0; j < n; ++j) {
            double ang = sign * 2.0 * M_PI * (double)((j * k) % n) / (double)n;
            double c = cos(ang), s = sin(ang);
            re += x[j].re * c - x[j].im * s;
            im += x[j].re * s + x[j].im * c;
        }
        y[k] = cx_make((REAL)re, (REAL)im);
    }
}

/* fftshift: out[k] = in[(k + d) mod n], d = n/2.
 * ifftshift: d = n/2 (+1 when n is odd).  include/fftutils.h:34-136 */
void or_shift(const REAL *in, REAL *out, int64_t n1, int64_t n2, int64_t n3,
              int axes, int inverse, int ncomp) {
    int64_t d3 = n3 / 2, d2 = n2 / 2;
    if (inverse) {
        if (n3 % 2) d3 += 1;
        if (n2 % 2) d2 += 1;
    }
    if (axes == 1) d2 = 0;
#pragma omp parallel for collapse(2)
    for (int64_t i = 0; i < n1; ++i)
        for (int64_t j = 0; j < n2; ++j) {
            int64_t j2 = (j + d2) % n2;
            for (int64_t k = 0; k < n3; ++k) {
                int64_t k2 = (k + d3) % n3;
                for (int c = 0; c < ncomp; ++c)
                    out[((i * n2 + j) * n3 + k) * ncomp + c] =
                        in[((i * n2 + j2) * n3 + k2) * ncomp + c];
            }
        }
}

int or_real_bytes(void) { return (int)sizeof(REAL); }
int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

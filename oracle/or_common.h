/* Shared internals of the CPU oracle (test infrastructure only). */
#ifndef OR_COMMON_H
#define OR_COMMON_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    REAL re, im;
} cx;

static inline cx cx_make(REAL a, REAL b) {
    cx z;
    z.re = a;
    z.im = b;
    return z;
}
static inline cx cx_add(cx a, cx b) { return cx_make(a.re + b.re, a.im + b.im); }
static inline cx cx_sub(cx a, cx b) { return cx_make(a.re - b.re, a.im - b.im); }
static inline cx cx_mul(cx a, cx b) {
    return cx_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cx_conj(cx a) { return cx_make(a.re, -a.im); }
static inline cx cx_scale(cx a, REAL s) { return cx_make(a.re * s, a.im * s); }
/* exp(i*ang) evaluated in double */
static inline cx cx_expi(double ang) {
    return cx_make((REAL)cos(ang), (REAL)sin(ang));
}

/* Twiddle-carrying FFT plan for one length. */
typedef struct or_fftplan {
    int64_t n;
    int nrad;
    int rad[64];
    cx *tw;           /* tw[k] = exp(-2 pi i k / n) */
    int blue;         /* 1: Bluestein through a power-of-two plan */
    int64_t m;
    cx *chirp;        /* exp(-i pi j^2 / n), j < n */
    cx *bfft;         /* FFT_m of the conjugate chirp sequence */
    struct or_fftplan *sub;
} or_fftplan;

or_fftplan *or_plan_get(int64_t n);   /* cached, thread safe */
/* transform one contiguous line; work must hold 2*n (4*m for Bluestein) cx */
void or_plan_exec(const or_fftplan *p, cx *line, cx *work, int sign);
size_t or_plan_work(const or_fftplan *p);

/* ES kernel helpers */
typedef struct {
    int w;
    double beta;
} or_es;
void or_es_choose(double tol, or_es *k);
double or_es_phi(const or_es *k, double z);           /* z in [-1,1] */
double or_es_phihat(const or_es *k, double xi);       /* int psi(u) e^{i xi u} du */
int64_t or_fine_size(int64_t nmodes, int w);

#endif

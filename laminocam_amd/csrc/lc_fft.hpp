// LDS-resident mixed-radix Stockham FFT core for gfx950.
//
// A workgroup transforms `nl` lines of length L held in LDS.  Two layouts:
//   ROW (LFAST=false): element e of line l at swz(l*ls + e)   (ls >= L)
//   COL (LFAST=true):  element e of line l at swz(e*nl + l)   (lines interleaved)
// swz() permutes addresses inside aligned 32-element groups so the strided
// Stockham writes spread over the 64 LDS banks (LDS buffers are allocated in
// multiples of 32 elements).  Each stage of radix R reads its R inputs into
// registers, synchronises, then writes the outputs in autosorted (Stockham)
// order back into the same buffer, so one LDS buffer suffices.
//
// Stage with radix R after Ns = product of the previous radices:
//   for j < L/R, k = j mod Ns:
//     v_r = in[j + r L/R] * w_L^{k r L/(Ns R)}      (w_L = e^{-2 pi i/L})
//     V   = DFT_R(v)
//     out[(j/Ns) Ns R + k + r Ns] = V_r
// Twiddles come from a table tw[k] = e^{-2 pi i k / L} generated in FP64.
#pragma once

#include <type_traits>

#include "lc_common.hpp"

namespace lc {

constexpr int FFT_MAX_STAGES = 16;

struct FftPlan {
    int L;
    int nst;
    int rad[FFT_MAX_STAGES];
};

// factor L = 2^a 3^b into radices 8/4/2/3; returns false if L has other factors
inline bool fft_make_plan(int L, FftPlan &p) {
    p.L = L;
    p.nst = 0;
    int n = L;
    int twos = 0, threes = 0;
    while (n % 2 == 0) { n /= 2; ++twos; }
    while (n % 3 == 0) { n /= 3; ++threes; }
    if (n != 1) return false;
    while (twos >= 3) { p.rad[p.nst++] = 8; twos -= 3; }
    if (twos == 2) p.rad[p.nst++] = 4;
    if (twos == 1) p.rad[p.nst++] = 2;
    for (int i = 0; i < threes; ++i) p.rad[p.nst++] = 3;
    return p.nst <= FFT_MAX_STAGES;
}

inline bool is_2_3_smooth(int64_t n) {
    if (n < 1) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

// LDS elements to allocate for `n` used elements (swz stays inside 32-groups)
__host__ __device__ constexpr int lds_round(int n) { return (n + 31) & ~31; }

__device__ __forceinline__ int swz(int a) { return a ^ (((a >> 5) & 7) << 2); }
__device__ __forceinline__ int lds_row(int line, int e, int ls) { return swz(line * ls + e); }
__device__ __forceinline__ int lds_col(int line, int e, int nl) { return swz(e * nl + line); }

// quotient of non-negative b < 2^22 by d >= 1 through an FP32 reciprocal,
// corrected to the exact value
__device__ __forceinline__ int fdiv(int b, int d, float inv) {
    int q = (int)((float)b * inv);
    int r = b - q * d;
    if (r < 0) { --q; r += d; }
    if (r >= d) ++q;
    return q;
}

// multiply by SIGN * i
template <int SIGN, typename C> __host__ __device__ __forceinline__ C mul_si(C a) {
    C r;
    if (SIGN < 0) { r.x = a.y; r.y = -a.x; }
    else { r.x = -a.y; r.y = a.x; }
    return r;
}

template <int SIGN, typename C> __host__ __device__ __forceinline__ void dft2(C *v) {
    C a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <int SIGN, typename C> __host__ __device__ __forceinline__ void dft4(C *v) {
    C a0 = v[0] + v[2], a1 = v[0] - v[2];
    C b0 = v[1] + v[3], b1 = v[1] - v[3];
    C jb1 = mul_si<SIGN>(b1);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + jb1;
    v[3] = a1 - jb1;
}

template <int SIGN, typename C> __host__ __device__ __forceinline__ void dft3(C *v) {
    using R = decltype(C{}.x);
    const R h = (R)0.8660254037844386467637231707529362;  // sqrt(3)/2
    C t1 = v[1] + v[2];
    C t2 = v[0] - cscale(t1, (R)0.5);
    C d = cscale(v[1] - v[2], h);
    C jd = mul_si<SIGN>(d);
    v[0] = v[0] + t1;
    v[1] = t2 + jd;
    v[2] = t2 - jd;
}

// 6 = 2 x 3: DFT3 of the even and of the odd samples, then one radix-2
// layer with the twiddles w^1 = (1/2, SIGN sqrt(3)/2), w^2 = (-1/2, SIGN sqrt(3)/2)
template <int SIGN, typename C> __host__ __device__ __forceinline__ void dft6(C *v) {
    using R = decltype(C{}.x);
    const R h = (R)0.5, s = (R)SIGN * (R)0.8660254037844386467637231707529362;
    C e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
    dft3<SIGN>(e);
    dft3<SIGN>(o);
    C t1, t2;
    t1.x = h * o[1].x - s * o[1].y;
    t1.y = h * o[1].y + s * o[1].x;
    t2.x = -h * o[2].x - s * o[2].y;
    t2.y = -h * o[2].y + s * o[2].x;
    v[0] = e[0] + o[0];
    v[3] = e[0] - o[0];
    v[1] = e[1] + t1;
    v[4] = e[1] - t1;
    v[2] = e[2] + t2;
    v[5] = e[2] - t2;
}

template <int SIGN, typename C> __host__ __device__ __forceinline__ void dft8(C *v) {
    using R = decltype(C{}.x);
    const R s = (R)0.7071067811865475244008443621048490;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4<SIGN>(e);
    dft4<SIGN>(o);
    // w = e^{SIGN 2 pi i / 8}
    C o1, o2, o3;
    o1.x = s * (o[1].x - SIGN * o[1].y);
    o1.y = s * (o[1].y + SIGN * o[1].x);
    o2 = mul_si<SIGN>(o[2]);
    o3.x = s * (-o[3].x - SIGN * o[3].y);
    o3.y = s * (-o[3].y + SIGN * o[3].x);
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
}

// 16 points as 4 x 4: DFT_4 over n1 of v[4 n1 + n2], twiddle w^{n2 k1}, DFT_4
// over n2; output k1 + 4 k2 (w = e^{SIGN 2 pi i / 16})
template <int SIGN, typename C> __host__ __device__ __forceinline__ void dft16(C *v) {
    using R = decltype(C{}.x);
    const R c1 = (R)0.9238795325112867561281831893967882;   // cos(pi/8)
    const R s1 = (R)0.3826834323650897717284599840303989;   // sin(pi/8)
    const R h = (R)0.7071067811865475244008443621048490;
    const C w1 = cmake<C>(c1, SIGN * s1), w2 = cmake<C>(h, SIGN * h), w3 = cmake<C>(s1, SIGN * c1);
    const C w6 = cmake<C>(-h, SIGN * h), w9 = cmake<C>(-c1, -SIGN * s1);
    C a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        a[n2][0] = v[n2];
        a[n2][1] = v[4 + n2];
        a[n2][2] = v[8 + n2];
        a[n2][3] = v[12 + n2];
        dft4<SIGN>(a[n2]);
    }
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = cmul(a[1][2], w2);
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2);
    a[2][2] = mul_si<SIGN>(a[2][2]);
    a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = cmul(a[3][2], w6);
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        C t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft4<SIGN>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}

template <int R, int SIGN, typename C> __host__ __device__ __forceinline__ void dftR(C *v) {
    if constexpr (R == 2) dft2<SIGN>(v);
    else if constexpr (R == 3) dft3<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v);
    else if constexpr (R == 6) dft6<SIGN>(v);
    else if constexpr (R == 8) dft8<SIGN>(v);
    else dft16<SIGN>(v);
}

template <typename T, int R, int SIGN, int NT, int MAXE, bool LFAST>
__device__ __forceinline__ void fft_stage(cpx<T> *buf, int nl, int L, int ls, int Ns,
                                          const cpx<T> *__restrict__ tw) {
    using C = cpx<T>;
    constexpr int MAXB = (MAXE + R - 1) / R;
    const int LR = L / R;
    const int nbf = nl * LR;
    const int twstep = L / (Ns * R);
    const int tid = threadIdx.x;
    const float inv_d = LFAST ? 1.0f / (float)nl : 1.0f / (float)LR;
    const float inv_ns = 1.0f / (float)Ns;
    C v[MAXB][R];
    C wa[MAXB], wb[MAXB];
    // twiddle loads first (unconditional: k = 0 reads w^0 = 1, an exact
    // multiply), so their latency overlaps the LDS reads and the barrier
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        int b = tid + t * NT;
        b = b < nbf ? b : nbf - 1;
        int j;
        if (LFAST) j = fdiv(b, nl, inv_d);
        else j = b - fdiv(b, LR, inv_d) * LR;
        const int k = j - fdiv(j, Ns, inv_ns) * Ns;
        wa[t] = tw[k * twstep];
        if constexpr (R == 8) wb[t] = tw[4 * k * twstep];
    }
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        const int b = tid + t * NT;
        if (b < nbf) {
            int line, j;
            if (LFAST) { j = fdiv(b, nl, inv_d); line = b - j * nl; }
            else { line = fdiv(b, LR, inv_d); j = b - line * LR; }
#pragma unroll
            for (int r = 0; r < R; ++r)
                v[t][r] = buf[LFAST ? lds_col(line, j + r * LR, nl) : lds_row(line, j + r * LR, ls)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        const int b = tid + t * NT;
        if (b < nbf) {
            int line, j;
            if (LFAST) { j = fdiv(b, nl, inv_d); line = b - j * nl; }
            else { line = fdiv(b, LR, inv_d); j = b - line * LR; }
            const int k = j - fdiv(j, Ns, inv_ns) * Ns;
            {
                // powers of w^k from one (radix 8: two) table entries
                C w[R];
                w[1] = wa[t];
                if (SIGN > 0) w[1].y = -w[1].y;
                if constexpr (R >= 3) w[2] = cmul(w[1], w[1]);
                if constexpr (R >= 4) w[3] = cmul(w[2], w[1]);
                if constexpr (R == 8) {
                    w[4] = wb[t];
                    if (SIGN > 0) w[4].y = -w[4].y;
                    w[5] = cmul(w[4], w[1]);
                    w[6] = cmul(w[4], w[2]);
                    w[7] = cmul(w[4], w[3]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], w[r]);
            }
            dftR<R, SIGN>(v[t]);
            const int o = (j - k) * R + k;   // (j/Ns)*Ns*R + k
#pragma unroll
            for (int r = 0; r < R; ++r)
                buf[LFAST ? lds_col(line, o + r * Ns, nl) : lds_row(line, o + r * Ns, ls)] = v[t][r];
        }
    }
    __syncthreads();
}

// Full transform of nl lines (all threads of the block must call this).
// Requires nl * L <= NT * MAXE.
template <typename T, int SIGN, int NT, int MAXE, bool LFAST>
__device__ void fft_lds(cpx<T> *buf, int nl, int ls, const FftPlan &plan,
                        const cpx<T> *__restrict__ tw) {
    int Ns = 1;
    const int L = plan.L;
    for (int s = 0; s < plan.nst; ++s) {
        const int R = plan.rad[s];
        switch (R) {
        case 8: fft_stage<T, 8, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        case 4: fft_stage<T, 4, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        case 3: fft_stage<T, 3, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        default: fft_stage<T, 2, SIGN, NT, MAXE, LFAST>(buf, nl, L, ls, Ns, tw); break;
        }
        Ns *= R;
    }
}

// ---------------------------------------------------------------------------
// Compile-time-length variant: the radix plan, strides and twiddle steps are
// constants, so the index arithmetic folds to shifts and constant multiplies
// and the stage loop unrolls.  Column layout (LFAST) requires the number of
// lines to be a power of two, passed as its log2.
// ---------------------------------------------------------------------------
constexpr int ct_twos(int L) { int t = 0; while (L % 2 == 0) { L /= 2; ++t; } return t; }
constexpr int ct_threes(int L) { int t = 0; while (L % 3 == 0) { L /= 3; ++t; } return t; }
// radices: as many RM (16 or 8) as divide L, one of 8 / 4 / 2 for the
// remaining power of two, and the 3s.  Two orders, chosen per length by
// timing the fused column convolution (fft_conv_ct, whose middle runs in the
// last stage) on gfx950:
//   RMs, remainder, 3s                      (L = 1024: 16 16 4; 1536: 8 8 8 3)
//   one RM, remainder, 3s, the other RMs    (L = 2048: 16 8 16; 768: 8 4 3 8)
constexpr int ct_lg(int RM) { return RM == 16 ? 4 : 3; }
constexpr bool ct_big_last(int L) { return L == 2048 || L == 768; }
// radix 16 with a remainder of one 2 and a single 3 (L = 2^(4a+1) 3, e.g.
// 1536 = 16 16 6): the 2 and the 3 merge into one radix-6 stage
constexpr bool ct_six(int L, int RM) {
    return RM == 16 && ct_threes(L) == 1 && ct_twos(L) % 4 == 1;
}
constexpr int ct_nst(int L, int RM = 8) {
    return ct_six(L, RM) ? ct_twos(L) / 4 + 1
                         : ct_twos(L) / ct_lg(RM) + (ct_twos(L) % ct_lg(RM) ? 1 : 0) + ct_threes(L);
}
constexpr int ct_radix(int L, int s, int RM = 8) {
    if (ct_six(L, RM)) return s < ct_twos(L) / 4 ? 16 : 6;
    const int nb = ct_twos(L) / ct_lg(RM), rem = ct_twos(L) % ct_lg(RM);
    const int nrem = rem ? 1 : 0, th = ct_threes(L);
    const int lead = ct_big_last(L) ? (nb > 0 ? 1 : 0) : nb;   // RMs before the remainder
    if (s < lead) return RM;
    s -= lead;
    if (s < nrem) return 1 << rem;
    s -= nrem;
    if (s < th) return 3;
    return RM;
}

// Swizzled LDS addresses of a thread's R elements x0 + r*S.  When the stride
// S is a multiple of 256 the swizzle (bits 5-7 folded into bits 2-4) commutes
// with it: swz(x0 + r*S) = swz(x0) + r*S, so one swizzle serves all R and the
// rest become constant offsets.
template <int S> __device__ __forceinline__ int swz_step(int sx0, int x0, int r) {
    if constexpr (S % 256 == 0) return sx0 + r * S;
    else return swz(x0 + r * S);
}

// LSH >= 0: the column layout's log2(lines) is a compile-time constant (all
// strides fold); LSH < 0: it is the runtime lsh.
// Pruning for zero-padded linear convolution (every line holds at most
// (L + 1) / 2 nonzero leading elements, and only that many leading outputs are
// used): PIN (first stage) treats inputs at index >= (L + 1) / 2 as zero
// without reading them; POUT (last stage) skips outputs at index >= (L + 1) / 2.
template <typename T, int L, int R, int NS, int SIGN, int NT, int MAXE, bool LFAST, int LSH,
          bool PIN = false, bool POUT = false>
__device__ __forceinline__ void fft_stage_ct(cpx<T> *buf, int nl, int lsh,
                                             const cpx<T> *__restrict__ tw) {
    // element j + r LR with j >= 0 is at least r LR: past the half when r LR >= (L + 1) / 2
    constexpr int HALF = (L + 1) / 2;
    using C = cpx<T>;
    constexpr int MAXB = (MAXE + R - 1) / R;
    constexpr int LR = L / R;
    constexpr int TWS = L / (NS * R);
    const int ls = LSH >= 0 ? LSH : lsh;
    const int nbf = LSH >= 0 ? (LR << (LSH >= 0 ? LSH : 0)) : nl * LR;
    const int tid = threadIdx.x;
    C v[MAXB][R];
    C wa[MAXB], wb[MAXB], wc[MAXB];
    // twiddle loads first (k = 0 reads w^0 = 1, an exact multiply) so their
    // latency overlaps the LDS reads and the barrier
    if constexpr (NS > 1) {
#pragma unroll
        for (int t = 0; t < MAXB; ++t) {
            int b = tid + t * NT;
            b = b < nbf ? b : nbf - 1;
            const int j = LFAST ? (b >> ls) : (b % LR);
            const int k = j % NS;
            wa[t] = tw[k * TWS];
            if constexpr (R >= 5) wb[t] = tw[4 * k * TWS];
            if constexpr (R == 16) wc[t] = tw[8 * k * TWS];
        }
    }
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        const int b = tid + t * NT;
        if (b < nbf) {
            if (LFAST) {
                // element j + r LR of line `line` sits at ((j + r LR) << ls) + line = b + r (LR << ls)
                if constexpr (LSH >= 0) {
                    constexpr int S = LR << (LSH >= 0 ? LSH : 0);
                    const int sb = swz(b);
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        v[t][r] = (PIN && r * LR >= HALF) ? cmake<C>(0, 0) : buf[swz_step<S>(sb, b, r)];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        v[t][r] = (PIN && r * LR >= HALF) ? cmake<C>(0, 0) : buf[swz(b + r * (LR << ls))];
                }
            } else {
                const int line = b / LR, j = b - line * LR;
                const int x0 = line * L + j, sx0 = swz(x0);
#pragma unroll
                for (int r = 0; r < R; ++r)
                    v[t][r] = (PIN && r * LR >= HALF) ? cmake<C>(0, 0) : buf[swz_step<LR>(sx0, x0, r)];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        const int b = tid + t * NT;
        if (b < nbf) {
            int j, line;
            if (LFAST) { j = b >> ls; line = b & ((1 << ls) - 1); }
            else { line = b / LR; j = b - line * LR; }
            const int k = j % NS;
            if constexpr (NS > 1) {
                C w[R];
                w[1] = wa[t];
                if (SIGN > 0) w[1].y = -w[1].y;
                if constexpr (R >= 3) w[2] = cmul(w[1], w[1]);
                if constexpr (R >= 4) w[3] = cmul(w[2], w[1]);
                if constexpr (R == 6) {
                    w[4] = wb[t];
                    if (SIGN > 0) w[4].y = -w[4].y;
                    w[5] = cmul(w[4], w[1]);
                }
                if constexpr (R >= 8) {
                    w[4] = wb[t];
                    if (SIGN > 0) w[4].y = -w[4].y;
                    w[5] = cmul(w[4], w[1]);
                    w[6] = cmul(w[4], w[2]);
                    w[7] = cmul(w[4], w[3]);
                }
                if constexpr (R == 16) {
                    // w^8 from the table, w^12 = w^8 w^4: every power within
                    // two products of a table entry
                    w[8] = wc[t];
                    if (SIGN > 0) w[8].y = -w[8].y;
                    w[9] = cmul(w[8], w[1]);
                    w[10] = cmul(w[8], w[2]);
                    w[11] = cmul(w[8], w[3]);
                    w[12] = cmul(w[8], w[4]);
                    w[13] = cmul(w[12], w[1]);
                    w[14] = cmul(w[12], w[2]);
                    w[15] = cmul(w[12], w[3]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], w[r]);
            }
            dftR<R, SIGN>(v[t]);
            const int o = (j - k) * R + k;
            if (LFAST) {
                const int ob = (o << ls) + line;
                // last stage (NS = L / R): output o + r NS = j + r LR, at least r LR
                if constexpr (LSH >= 0) {
                    constexpr int S = NS << (LSH >= 0 ? LSH : 0);
                    const int sb = swz(ob);
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (!(POUT && r * LR >= HALF)) buf[swz_step<S>(sb, ob, r)] = v[t][r];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (!(POUT && r * LR >= HALF)) buf[swz(ob + r * (NS << ls))] = v[t][r];
                }
            } else {
                const int ob = line * L + o, sb = swz(ob);
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (!(POUT && r * LR >= HALF)) buf[swz_step<NS>(sb, ob, r)] = v[t][r];
            }
        }
    }
    __syncthreads();
}

// Radix-16 stages for FP32 power-of-two lengths >= 1024 when every thread
// holds at least 16 elements: one stage (an LDS round trip and two barriers)
// fewer than radix 8.  Elsewhere radix 8: the stage count is the same or
// the larger radix measured slower (lower occupancy).
template <typename T, int MAXE, int L> constexpr int ct_rmax() {
    return (sizeof(T) == 4 && MAXE >= 16 && L >= 1024 && (ct_threes(L) == 0 || ct_six(L, 16))) ? 16 : 8;
}

// Stages S .. SEND - 1 of the plan (REV: the radices in reverse order).  PIN
// prunes the plan's first stage, POUT its last.
template <typename T, int L, int S, int NS, int SIGN, int NT, int MAXE, bool LFAST, int LSH,
          bool PIN, bool POUT, bool REV = false, int SEND = -1>
__device__ __forceinline__ void fft_stages_ct(cpx<T> *buf, int nl, int lsh,
                                              const cpx<T> *__restrict__ tw) {
    constexpr int RM = ct_rmax<T, MAXE, L>();
    constexpr int NST = ct_nst(L, RM);
    constexpr int END = SEND < 0 ? NST : SEND;
    constexpr int R = ct_radix(L, REV ? NST - 1 - S : S, RM);
    constexpr bool LAST = S + 1 == END;
    fft_stage_ct<T, L, R, NS, SIGN, NT, MAXE, LFAST, LSH, PIN && S == 0, POUT && S + 1 == NST>(
        buf, nl, lsh, tw);
    if constexpr (!LAST)
        fft_stages_ct<T, L, S + 1, NS * R, SIGN, NT, MAXE, LFAST, LSH, PIN, POUT, REV, SEND>(buf, nl,
                                                                                           lsh, tw);
}

// Full transform of nl lines of compile-time length L (row layout: line
// stride L; column layout: nl = 1 << lsh lines interleaved, lsh = LSH when
// LSH >= 0).  PIN / POUT: the zero-padded-input / leading-output pruning of
// fft_stage_ct.
template <typename T, int L, int SIGN, int NT, int MAXE, bool LFAST, int LSH = -1,
          bool PIN = false, bool POUT = false>
__device__ __forceinline__ void fft_lds_ct(cpx<T> *buf, int nl, int lsh,
                                           const cpx<T> *__restrict__ tw) {
    fft_stages_ct<T, L, 0, 1, SIGN, NT, MAXE, LFAST, LSH, PIN, POUT>(buf, nl, lsh, tw);
}

// Line lengths with compiled (unrolled) line-pass FFTs (fft_lds_ct) in the
// NUFFT projectors: f(std::integral_constant<int, L>) for those, false for
// any other length (the runtime plan).
template <typename F> bool with_ct_len(int L, F &&f) {
    switch (L) {
    case 16: f(std::integral_constant<int, 16>{}); return true;
    case 32: f(std::integral_constant<int, 32>{}); return true;
    case 64: f(std::integral_constant<int, 64>{}); return true;
    case 96: f(std::integral_constant<int, 96>{}); return true;
    case 128: f(std::integral_constant<int, 128>{}); return true;
    case 192: f(std::integral_constant<int, 192>{}); return true;
    case 256: f(std::integral_constant<int, 256>{}); return true;
    case 384: f(std::integral_constant<int, 384>{}); return true;
    case 512: f(std::integral_constant<int, 512>{}); return true;
    case 768: f(std::integral_constant<int, 768>{}); return true;
    case 1024: f(std::integral_constant<int, 1024>{}); return true;
    case 1458: f(std::integral_constant<int, 1458>{}); return true;   // 2 * 721 (2 3^6)
    case 1536: f(std::integral_constant<int, 1536>{}); return true;
    case 2048: f(std::integral_constant<int, 2048>{}); return true;
    default: return false;
    }
}

// ---------------------------------------------------------------------------
// Column-layout circular convolution with a real spectrum:
//   line <- IFFT(kr .* FFT(line))   (unnormalised inverse)
// The forward transform's last stage leaves butterfly j holding outputs
// j + r L/R (r < R) in registers; the inverse transform, with its radices in
// reverse order, starts with a radix-R stage over those same indices.  So the
// two stages and the spectrum multiply run in registers: no LDS round trip
// between the transforms.  kr[t][r] is the spectrum at index j + r L/R of
// butterfly t of this thread (see conv_mid_index).  The forward input is
// pruned (PIN) and the inverse output (POUT) as in fft_stage_ct.
// ---------------------------------------------------------------------------
template <typename T, int L, int MAXE> struct ConvMid {
    static constexpr int RM = ct_rmax<T, MAXE, L>();
    static constexpr int NST = ct_nst(L, RM);
    static constexpr int R = ct_radix(L, NST - 1, RM);   // forward last = inverse first
    static constexpr int LR = L / R;
    static constexpr int MAXB = (MAXE + R - 1) / R;
};

// line and index j of this thread's mid-stage butterfly t (false if none)
template <typename T, int L, int NT, int MAXE, int LSH>
__device__ __forceinline__ bool conv_mid_index(int t, int nl, int lsh, int &line, int &j) {
    constexpr int LR = ConvMid<T, L, MAXE>::LR;
    const int ls = LSH >= 0 ? LSH : lsh;
    const int b = threadIdx.x + t * NT;
    if (b >= nl * LR) return false;
    j = b >> ls;
    line = b & ((1 << ls) - 1);
    return true;
}

template <typename T, int L, int NT, int MAXE, int LSH>
__device__ __forceinline__ void fft_conv_mid(cpx<T> *buf, int nl, int lsh,
                                             const cpx<T> *__restrict__ tw,
                                             const T (&kr)[ConvMid<T, L, MAXE>::MAXB]
                                                          [ConvMid<T, L, MAXE>::R]) {
    using C = cpx<T>;
    using CM = ConvMid<T, L, MAXE>;
    constexpr int R = CM::R, LR = CM::LR, MAXB = CM::MAXB;
    const int ls = LSH >= 0 ? LSH : lsh;
    C v[MAXB][R];
    C wa[MAXB], wb[MAXB], wc[MAXB];
    // forward last stage: NS = L / R, so k = j and the twiddle step is 1
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        int line, j;
        if (!conv_mid_index<T, L, NT, MAXE, LSH>(t, nl, lsh, line, j)) j = 0;
        wa[t] = tw[j];
        if constexpr (R >= 5) wb[t] = tw[4 * j];
        if constexpr (R == 16) wc[t] = tw[8 * j];
    }
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        int line, j;
        if (conv_mid_index<T, L, NT, MAXE, LSH>(t, nl, lsh, line, j)) {
            const int b = (j << ls) + line;
#pragma unroll
            for (int r = 0; r < R; ++r) v[t][r] = buf[swz(b + r * (LR << ls))];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        int line, j;
        if (conv_mid_index<T, L, NT, MAXE, LSH>(t, nl, lsh, line, j)) {
            C w[R];
            w[1] = wa[t];
            if constexpr (R >= 3) w[2] = cmul(w[1], w[1]);
            if constexpr (R >= 4) w[3] = cmul(w[2], w[1]);
            if constexpr (R == 6) {
                w[4] = wb[t];
                w[5] = cmul(w[4], w[1]);
            }
            if constexpr (R >= 8) {
                w[4] = wb[t];
                w[5] = cmul(w[4], w[1]);
                w[6] = cmul(w[4], w[2]);
                w[7] = cmul(w[4], w[3]);
            }
            if constexpr (R == 16) {
                w[8] = wc[t];
                w[9] = cmul(w[8], w[1]);
                w[10] = cmul(w[8], w[2]);
                w[11] = cmul(w[8], w[3]);
                w[12] = cmul(w[8], w[4]);
                w[13] = cmul(w[12], w[1]);
                w[14] = cmul(w[12], w[2]);
                w[15] = cmul(w[12], w[3]);
            }
#pragma unroll
            for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], w[r]);
            dftR<R, -1>(v[t]);
#pragma unroll
            for (int r = 0; r < R; ++r) v[t][r] = cscale(v[t][r], kr[t][r]);
            // inverse first stage (NS = 1, no twiddles): outputs j R + r
            dftR<R, 1>(v[t]);
            const int ob = ((j * R) << ls) + line;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[swz(ob + (r << ls))] = v[t][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int NT, int MAXE, int LSH>
__device__ __forceinline__ void fft_conv_ct(cpx<T> *buf, int nl, int lsh,
                                            const cpx<T> *__restrict__ tw,
                                            const T (&kr)[ConvMid<T, L, MAXE>::MAXB]
                                                         [ConvMid<T, L, MAXE>::R]) {
    using CM = ConvMid<T, L, MAXE>;
    static_assert(CM::NST >= 2, "fused convolution needs at least two stages");
    fft_stages_ct<T, L, 0, 1, -1, NT, MAXE, true, LSH, true, false, false, CM::NST - 1>(buf, nl, lsh,
                                                                                      tw);
    fft_conv_mid<T, L, NT, MAXE, LSH>(buf, nl, lsh, tw, kr);
    fft_stages_ct<T, L, 1, CM::R, 1, NT, MAXE, true, LSH, false, true, true>(buf, nl, lsh, tw);
}

// FP64 twiddle generation: tw[k] = exp(-2 pi i k / L)
template <typename T>
__global__ void k_twiddles(cpx<T> *tw, int L) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L) {
        double s, c;
        sincospi(-2.0 * (double)k / (double)L, &s, &c);
        cpx<T> z;
        z.x = (T)c;
        z.y = (T)s;
        tw[k] = z;
    }
}

} // namespace lc

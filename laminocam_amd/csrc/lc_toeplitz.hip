// 2-D Toeplitz normal operator A^T A on gfx950.
//
// Reference semantics (include/toeplitz2d.h:101-125): y = crop_{M/2}(C2R(
// R2C(pad_right(x)) * khat)) / (M^2 N), with khat = R2C of the CMCL-ordered
// PSF.  The PSF P(m) is real and even, so the spectrum of the *centred* PSF is
// real; using it and cropping at offset 0 is the same linear convolution.
// 1/(M^2 N) is folded into the stored spectrum.
//
// The apply is three HBM passes over each slice:
//   P1 rows:    pairs of real rows packed as one complex row, FFT_M, split
//               into two half spectra -> U
//   P2 columns: per column kx < H: load N values, FFT_M, x khat[kx][:],
//               IFFT_M, keep N values -> U in place
//   P3 rows:    pairs of half spectra rebuilt into one complex row by
//               Hermitian symmetry, IFFT_M, real/imag -> the two output rows
// U is stored column-blocked, [ceil(H/B)][N][B] complex, so that P2 reads and
// writes one contiguous N x B block and P1/P3 touch 2*ppb*B contiguous
// elements per block.
// Algorithmic bytes per slice: 8 N^2 + 32 N H + 4 M H.
#include "lc_engine.hpp"

namespace lc {

extern __shared__ __align__(16) char lc_smem[];

int64_t next_fast_dim(int64_t n) {
    int64_t best = 1;
    while (best < n) best <<= 1;
    for (int64_t q3 = 3; q3 < best; q3 *= 3) {
        int64_t x = q3;
        while (x < n) x <<= 1;
        if (x < best) best = x;
    }
    return best;
}

// ----------------------------------------------------------------------------
// FP64 PSF: P(m) = sum_i D_nrad(m_x cos th_i + m_y sin th_i),
// D_n(t) = sin(pi t)/sin(pi t/n) (the radial sum of exp(i r_k t) over the
// polar grid's r_k = (k+1/2) 2pi/n - pi), evaluated at t = k n + d as
// (-1)^{k(n-1)} sin(pi d)/sin(pi d/n) so it never divides 0 by 0.
// ----------------------------------------------------------------------------
__device__ __forceinline__ double dirichlet_d(double t, int n, double inv_n) {
    const double kq = rint(t * inv_n);
    const double d = fma(-kq, (double)n, t);
    const long long k = (long long)kq;
    const double sg = ((k * (long long)(n - 1)) & 1LL) ? -1.0 : 1.0;
    if (d == 0.0) return sg * (double)n;
    return sg * sinpi(d) / sinpi(d * inv_n);
}

constexpr int PSF_RUN = 8;          // consecutive kx per thread (rotation recurrence)
constexpr int PSF_ANG_CHUNK = 512;  // angles staged in LDS per pass

// P in periodic (centred) layout: index j <-> m = j (j < M/2) or j - M.
// Rows iy <= M/2 are computed and mirrored through P(-m) = P(m).  Along a run
// of consecutive kx the phases exp(i pi t) and exp(i pi t / n) advance by a
// fixed rotation per step; the ratio of their imaginary parts is D_n(t).
// Where sin(pi t / n) is small the direct reduced form is used instead.
__global__ void __launch_bounds__(256)
k_psf2d_gen(double2 *P, const double *theta, int nang, int nrad, int M) {
    __shared__ double4 ang[PSF_ANG_CHUNK];    // (cos, sin, cos pi c, sin pi c)
    __shared__ double2 ang2[PSF_ANG_CHUNK];   // (cos pi c/n, sin pi c/n)
    const double inv_n = 1.0 / (double)nrad;
    const int runs = (M + PSF_RUN - 1) / PSF_RUN;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = gid / runs;
    const int run = gid - iy * runs;
    const bool valid = iy <= M / 2;      // no early return: the block synchronises below
    const int ix0 = run * PSF_RUN;
    const int half = M / 2;
    const double ky = (double)(iy < (M + 1) / 2 ? iy : iy - M);
    double acc[PSF_RUN];
#pragma unroll
    for (int j = 0; j < PSF_RUN; ++j) acc[j] = 0.0;
    for (int a0 = 0; a0 < nang; a0 += PSF_ANG_CHUNK) {
        const int na = min(PSF_ANG_CHUNK, nang - a0);
        __syncthreads();
        for (int i = threadIdx.x; i < na; i += blockDim.x) {
            double s, c;
            sincos(theta[a0 + i], &s, &c);
            double s1, c1, s2, c2;
            sincospi(c, &s1, &c1);
            sincospi(c * inv_n, &s2, &c2);
            ang[i] = make_double4(c, s, c1, s1);
            ang2[i] = make_double2(c2, s2);
        }
        __syncthreads();
        if (!valid) continue;
        for (int i = 0; i < na; ++i) {
            const double4 a = ang[i];
            const double2 a2 = ang2[i];
            double e1r = 0, e1i = 0, e2r = 0, e2i = 0;
#pragma unroll
            for (int j = 0; j < PSF_RUN; ++j) {
                const int ix = ix0 + j;
                if (ix >= M) break;
                const double kx = (double)(ix < (M + 1) / 2 ? ix : ix - M);
                const double t = kx * a.x + ky * a.y;
                if (j == 0 || ix == (M + 1) / 2) {
                    sincospi(t, &e1i, &e1r);
                    sincospi(t * inv_n, &e2i, &e2r);
                } else {
                    const double r1 = e1r * a.z - e1i * a.w, i1 = e1r * a.w + e1i * a.z;
                    const double r2 = e2r * a2.x - e2i * a2.y, i2 = e2r * a2.y + e2i * a2.x;
                    e1r = r1; e1i = i1; e2r = r2; e2i = i2;
                }
                acc[j] += (fabs(e2i) > 1e-3) ? e1i / e2i : dirichlet_d(t, nrad, inv_n);
            }
        }
    }
    if (!valid) return;
    // rows 0 and M/2 (even M) are computed in full and own themselves; rows
    // strictly inside also fill their mirror row, so every entry has exactly
    // one writer
    const bool mirror = iy >= 1 && 2 * iy < M;
    (void)half;
#pragma unroll
    for (int j = 0; j < PSF_RUN; ++j) {
        const int ix = ix0 + j;
        if (ix >= M) break;
        P[(size_t)iy * M + ix] = make_double2(acc[j], 0.0);
        if (mirror) P[(size_t)(M - iy) * M + (M - ix) % M] = make_double2(acc[j], 0.0);
    }
}

// in-place complex FFT of `nrows` contiguous lines of length L; nl lines per block
template <typename T, int SIGN, int NT, int MAXE>
__global__ void __launch_bounds__(NT) k_rows_c2c(cpx<T> *data, int nrows, int nl, FftPlan plan,
                                                 const cpx<T> *__restrict__ tw) {
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(lc_smem);
    const int L = plan.L;
    const int row0 = blockIdx.x * nl;
    const int nloc = min(nl, nrows - row0);
    for (int idx = threadIdx.x; idx < nl * L; idx += NT) {
        const int l = idx / L, e = idx - l * L;
        buf[lds_row(l, e, L)] = (l < nloc) ? data[(size_t)(row0 + l) * L + e] : cpx<T>{};
    }
    __syncthreads();
    fft_lds<T, SIGN, NT, MAXE, false>(buf, nl, L, plan, tw);
    for (int idx = threadIdx.x; idx < nloc * L; idx += NT) {
        const int l = idx / L, e = idx - l * L;
        data[(size_t)(row0 + l) * L + e] = buf[lds_row(l, e, L)];
    }
}

// FP64 column FFT of the row-transformed PSF, keeping columns kx < H:
// khatT[kx][ky] = Re(FFT)[ky][kx] * scale (float)
template <int NT, int MAXE>
__global__ void __launch_bounds__(NT) k_psf_cols(const double2 *P, float *khatT, int M, int H, int B,
                                                 FftPlan plan, const double2 *__restrict__ tw,
                                                 double scale) {
    double2 *buf = reinterpret_cast<double2 *>(lc_smem);
    const int col0 = blockIdx.x * B;
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        buf[swz(idx)] = (col0 + c < H) ? P[(size_t)i * M + col0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, -1, NT, MAXE, true>(buf, B, 0, plan, tw);
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int c = idx / M, i = idx - c * M;
        if (col0 + c < H) khatT[(size_t)(col0 + c) * M + i] = (float)(buf[lds_col(c, i, B)].x * scale);
    }
}

// reference-layout kernel_hat from the stored real spectrum.  The reference
// transforms the PSF stored with its centre at index h = floor(M/2), which
// multiplies the centred spectrum by exp(-2 pi i (kx + ky) h / M):
// khat_ref[ky][kx] = khatT[kx][ky] * exp(-2 pi i (kx+ky) h / M) * (M^2 N)
__global__ void k_khat_ref(const float *khatT, float2 *out, int M, int H, float unscale) {
    const int kx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ky = blockIdx.y;
    if (kx >= H) return;
    const float v = khatT[(size_t)kx * M + ky] * unscale;
    const long long ph = ((long long)(kx + ky) * (M / 2)) % M;
    double s, c;
    sincospi(-2.0 * (double)ph / (double)M, &s, &c);
    out[(size_t)ky * H + kx] = make_float2((float)(v * c), (float)(v * s));
}

void geom2d_init(Ctx &c, Geom2D &g, const double *theta, int64_t nang, int64_t nrad,
                 int64_t image_n) {
    LC_REQUIRE(nang >= 1 && nrad >= 2 && image_n >= 2, "geometry: need nang>=1, nrad>=2, N>=2");
    LC_REQUIRE(nang <= (1 << 20), "geometry: too many angles");
    g.nang = nang;
    g.nrad = nrad;
    g.N = image_n;
    g.M = next_fast_dim(2 * image_n - 1);
    g.H = g.M / 2 + 1;
    LC_REQUIRE(g.M <= 4096, "geometry: image size above 2048 is not supported");
    g.theta.assign(theta, theta + nang);
    g.d_theta.alloc(sizeof(double) * nang);
    LC_HIP(hipMemcpyAsync(g.d_theta.as(), theta, sizeof(double) * nang, hipMemcpyHostToDevice,
                          c.stream()));
    c.sync();
    g.psf_ready = false;
}

void geom2d_build_psf(Ctx &c, Geom2D &g) {
    const int M = (int)g.M, H = (int)g.H;
    FftPlan plan;
    LC_REQUIRE(fft_make_plan(M, plan), "PSF size is not 2^p 3^q");
    double2 *P = reinterpret_cast<double2 *>(c.scratch(0, sizeof(double2) * (size_t)M * M));
    const int runs = (M + PSF_RUN - 1) / PSF_RUN;
    const int threads = (M / 2 + 1) * runs;
    hipLaunchKernelGGL(k_psf2d_gen, dim3(ceil_div(threads, 256)), dim3(256), 0, c.stream(), P,
                       g.d_theta.as<double>(), (int)g.nang, (int)g.nrad, M);
    LC_HIP(hipGetLastError());
    const double2 *tw = c.twiddles<double>(M);
    constexpr int NT = 256, MAXE = 16;
    // rows: nl lines per block with nl*M <= NT*MAXE
    const int nl = std::max(1, NT * MAXE / M);
    hipLaunchKernelGGL((k_rows_c2c<double, -1, NT, MAXE>), dim3(ceil_div(M, nl)), dim3(NT),
                       sizeof(double2) * lds_round(nl * M), c.stream(), P, M, nl, plan, tw);
    LC_HIP(hipGetLastError());
    g.khatT.alloc(sizeof(float) * (size_t)H * M);
    const int B = std::max(1, NT * MAXE / M);
    const double scale = 1.0 / ((double)M * (double)M * (double)g.N);
    hipLaunchKernelGGL((k_psf_cols<NT, MAXE>), dim3(ceil_div(H, B)), dim3(NT),
                       sizeof(double2) * lds_round(B * M), c.stream(), P, g.khatT.as<float>(), M,
                       H, B, plan, tw, scale);
    LC_HIP(hipGetLastError());
    g.psf_ready = true;
}

void geom2d_kernel_hat(Ctx &c, Geom2D &g, float *out_host) {
    const int M = (int)g.M, H = (int)g.H;
    float2 *tmp = reinterpret_cast<float2 *>(c.scratch(1, sizeof(float2) * (size_t)M * H));
    const float unscale = (float)((double)M * (double)M * (double)g.N);
    hipLaunchKernelGGL(k_khat_ref, dim3(ceil_div(H, 256), M), dim3(256), 0, c.stream(),
                       g.khatT.as<float>(), tmp, M, H, unscale);
    LC_HIP(hipGetLastError());
    LC_HIP(hipMemcpyAsync(out_host, tmp, sizeof(float2) * (size_t)M * H, hipMemcpyDeviceToHost,
                          c.stream()));
    c.sync();
}

// ----------------------------------------------------------------------------
// Toeplitz passes
// ----------------------------------------------------------------------------

__device__ __forceinline__ bool prob_active(const CgState *st, int prob) {
    return st == nullptr || st[prob].active != 0;
}

// P1: U from pairs of real rows.  FUSE: the row source is the CG direction
// p = r + beta p (written back) instead of a plain input.
template <int NT, int MAXE, bool FUSE>
__global__ void __launch_bounds__(NT)
k_toep_rows_fwd(const SliceRef *__restrict__ sl, const CgState *__restrict__ st,
                const float *src, const float *__restrict__ rsrc, float *pdst,
                float2 *__restrict__ U, ToepGeom g, FftPlan plan, const float2 *__restrict__ tw) {
    const SliceRef s = sl[blockIdx.y];
    if (!prob_active(st, s.prob)) return;
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    const int N = g.N, M = g.M, H = g.H, B = g.B, ppb = g.ppb;
    const int a0 = 2 * blockIdx.x * ppb;
    const int nr = min(2 * ppb, N - a0);
    double beta = 0.0;
    if (FUSE) beta = st[s.prob].beta;
    const float *x = src + s.xoff + (int64_t)a0 * N;
    const float *rr = FUSE ? rsrc + s.xoff + (int64_t)a0 * N : nullptr;
    float *pw = FUSE ? pdst + s.xoff + (int64_t)a0 * N : nullptr;
    const int half = M / 2;
    if (((N | M) & 1) == 0) {
        // items: (line l, column pair q) over ppb x M/2; loads batched first
        constexpr int LPT = MAXE / 2;
        const int nitems = ppb * half;
        const float inv = 1.0f / (float)half;
        for (int base = 0; base < nitems; base += NT * LPT) {
            float2 va[LPT], vb[LPT];
#pragma unroll
            for (int t = 0; t < LPT; ++t) {
                const int it = base + threadIdx.x + t * NT;
                va[t] = make_float2(0.f, 0.f);
                vb[t] = make_float2(0.f, 0.f);
                if (it < nitems) {
                    const int l = fdiv(it, half, inv), q = it - l * half;
                    const int n = 2 * q;
                    if (n < N) {
                        const int r0 = 2 * l;
                        if (r0 < nr) va[t] = *reinterpret_cast<const float2 *>(x + (int64_t)r0 * N + n);
                        if (r0 + 1 < nr) vb[t] = *reinterpret_cast<const float2 *>(x + (int64_t)(r0 + 1) * N + n);
                        if (FUSE) {
                            float2 ra = make_float2(0.f, 0.f), rb = ra;
                            if (r0 < nr) ra = *reinterpret_cast<const float2 *>(rr + (int64_t)r0 * N + n);
                            if (r0 + 1 < nr) rb = *reinterpret_cast<const float2 *>(rr + (int64_t)(r0 + 1) * N + n);
                            va[t] = make_float2((float)(ra.x + beta * va[t].x), (float)(ra.y + beta * va[t].y));
                            vb[t] = make_float2((float)(rb.x + beta * vb[t].x), (float)(rb.y + beta * vb[t].y));
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < LPT; ++t) {
                const int it = base + threadIdx.x + t * NT;
                if (it < nitems) {
                    const int l = fdiv(it, half, inv), q = it - l * half;
                    const int n = 2 * q;
                    if (FUSE && n < N) {
                        const int r0 = 2 * l;
                        if (r0 < nr) *reinterpret_cast<float2 *>(pw + (int64_t)r0 * N + n) = va[t];
                        if (r0 + 1 < nr) *reinterpret_cast<float2 *>(pw + (int64_t)(r0 + 1) * N + n) = vb[t];
                    }
                    buf[lds_row(l, n, M)] = make_float2(va[t].x, vb[t].x);
                    buf[lds_row(l, n + 1, M)] = make_float2(va[t].y, vb[t].y);
                }
            }
        }
    } else {
        for (int it = threadIdx.x; it < ppb * M; it += NT) {
            const int l = it / M, n = it - l * M;
            float re = 0.f, im = 0.f;
            if (n < N) {
                const int r0 = 2 * l;
                if (r0 < nr) {
                    re = x[(int64_t)r0 * N + n];
                    if (FUSE) { re = (float)(rr[(int64_t)r0 * N + n] + beta * re); pw[(int64_t)r0 * N + n] = re; }
                }
                if (r0 + 1 < nr) {
                    im = x[(int64_t)(r0 + 1) * N + n];
                    if (FUSE) { im = (float)(rr[(int64_t)(r0 + 1) * N + n] + beta * im); pw[(int64_t)(r0 + 1) * N + n] = im; }
                }
            }
            buf[lds_row(l, n, M)] = make_float2(re, im);
        }
    }
    __syncthreads();
    fft_lds<float, -1, NT, MAXE, false>(buf, ppb, M, plan, tw);
    // split into half spectra and store column-blocked: items (kb, row, c)
    float2 *u = U + s.uoff;
    const int per_blk = nr * B;
    const int nitems = g.nblk * per_blk;
    const float inv_pb = 1.0f / (float)per_blk, inv_b = 1.0f / (float)B;
    for (int it = threadIdx.x; it < nitems; it += NT) {
        const int kb = fdiv(it, per_blk, inv_pb);
        const int rem = it - kb * per_blk;
        const int row = fdiv(rem, B, inv_b), c = rem - row * B;
        const int k = kb * B + c;
        float2 v = make_float2(0.f, 0.f);
        if (k < H) {
            const int l = row >> 1;
            const float2 z = buf[lds_row(l, k, M)];
            const float2 zc = cconj(buf[lds_row(l, k == 0 ? 0 : M - k, M)]);
            // X_a = (Z + conj Z(-k))/2,  X_b = (Z - conj Z(-k))/(2i)
            v = (row & 1) ? make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x))
                          : make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        }
        u[((int64_t)kb * N + a0 + row) * B + c] = v;
    }
}

// P2: column FFT, multiply by the real spectrum, inverse FFT, crop to N rows.
// The spectrum columns are fetched into registers before the first FFT so
// their latency hides behind it.
template <int NT, int MAXE, bool PF>
__global__ void __launch_bounds__(NT)
k_toep_cols(const SliceRef *__restrict__ sl, const ProbDev *__restrict__ pd,
            const CgState *__restrict__ st, float2 *__restrict__ U, ToepGeom g, FftPlan plan,
            const float2 *__restrict__ tw) {
    const int nbx = g.nblk;
    const int lin = blockIdx.x + blockIdx.y * gridDim.x;
    const int logical = xcd_remap(lin, gridDim.x * gridDim.y);
    const int sidx = logical / nbx, kb = logical - sidx * nbx;
    const SliceRef s = sl[sidx];
    if (!prob_active(st, s.prob)) return;
    const int N = g.N, M = g.M, H = g.H, B = g.B;
    const float *__restrict__ kh = pd[s.prob].khatT;
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    float4 *blk4 = reinterpret_cast<float4 *>(U + s.uoff + (int64_t)kb * N * B);
    const int nval = N * B;            // complex values in the block (kept even)
    const int tot = M * B;             // <= NT * MAXE
    const bool vec_k = (M & 3) == 0;
    // spectrum columns: B*M/4 float4 <= NT*MAXE/4 (fetched before the first
    // FFT when PF, after it otherwise)
    constexpr int KPT = MAXE / 4;
    float4 kv[KPT];
    const int m4 = M >> 2;
    const float inv_m4 = 1.0f / (float)(m4 > 0 ? m4 : 1);
    auto fetch_k = [&]() {
#pragma unroll
        for (int t = 0; t < KPT; ++t) {
            const int it = threadIdx.x + t * NT;
            kv[t] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (it < B * m4) {
                const int c = fdiv(it, m4, inv_m4), i4 = it - c * m4;
                if (kb * B + c < H)
                    kv[t] = *reinterpret_cast<const float4 *>(kh + (int64_t)(kb * B + c) * M + 4 * i4);
            }
        }
    };
    if (PF && vec_k) fetch_k();
    // block load: (tot+1)/2 float4 <= NT*MAXE/2; rows >= N are zero
    {
        constexpr int LPT = MAXE / 4;     // two passes of MAXE/4 float4
        const int n4 = nval >> 1;
        const int t2 = (tot + 1) >> 1;
        for (int base = 0; base < t2; base += NT * LPT) {
            float4 v[LPT];
#pragma unroll
            for (int t = 0; t < LPT; ++t) {
                const int it = base + threadIdx.x + t * NT;
                v[t] = (it < n4) ? blk4[it] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int t = 0; t < LPT; ++t) {
                const int it = base + threadIdx.x + t * NT;
                if (it < t2) {
                    buf[swz(2 * it)] = make_float2(v[t].x, v[t].y);
                    if (2 * it + 1 < tot) buf[swz(2 * it + 1)] = make_float2(v[t].z, v[t].w);
                }
            }
        }
    }
    __syncthreads();
    fft_lds<float, -1, NT, MAXE, true>(buf, B, 0, plan, tw);
    if (vec_k) {
        if (!PF) fetch_k();
#pragma unroll
        for (int t = 0; t < KPT; ++t) {
            const int it = threadIdx.x + t * NT;
            if (it < B * m4) {
                const int c = fdiv(it, m4, inv_m4), i = 4 * (it - c * m4);
                float2 *p0 = buf + lds_col(c, i, B);
                *p0 = cscale(*p0, kv[t].x);
                float2 *p1 = buf + lds_col(c, i + 1, B);
                *p1 = cscale(*p1, kv[t].y);
                float2 *p2 = buf + lds_col(c, i + 2, B);
                *p2 = cscale(*p2, kv[t].z);
                float2 *p3 = buf + lds_col(c, i + 3, B);
                *p3 = cscale(*p3, kv[t].w);
            }
        }
    } else {
        for (int it = threadIdx.x; it < B * M; it += NT) {
            const int c = it / M, i = it - c * M;
            const float k = (kb * B + c < H) ? kh[(int64_t)(kb * B + c) * M + i] : 0.0f;
            float2 *p0 = buf + lds_col(c, i, B);
            *p0 = cscale(*p0, k);
        }
    }
    __syncthreads();
    fft_lds<float, 1, NT, MAXE, true>(buf, B, 0, plan, tw);
    {
        const int n4 = nval >> 1;
        for (int it = threadIdx.x; it < n4; it += NT) {
            const float2 a = buf[swz(2 * it)], b = buf[swz(2 * it + 1)];
            blk4[it] = make_float4(a.x, a.y, b.x, b.y);
        }
    }
}

// P3: rebuild full spectra of two rows, inverse FFT, write real rows.
// DOT: also the block's partial of <p, Ap> into partial[slot].
template <int NT, int MAXE, bool DOT>
__global__ void __launch_bounds__(NT)
k_toep_rows_inv(const SliceRef *__restrict__ sl, const CgState *__restrict__ st,
                const float2 *__restrict__ U, float *__restrict__ dst,
                const float *__restrict__ pvec, double *__restrict__ partial, int slot_base,
                ToepGeom g, FftPlan plan, const float2 *__restrict__ tw) {
    const SliceRef s = sl[blockIdx.y];
    if (!prob_active(st, s.prob)) return;
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    const int N = g.N, M = g.M, H = g.H, B = g.B, ppb = g.ppb;
    const int a0 = 2 * blockIdx.x * ppb;
    const int nr = min(2 * ppb, N - a0);
    const int np = (nr + 1) >> 1;
    const float2 *u = U + s.uoff;
    const int half = M / 2;
    // gather: items (kb, l, c); each reads rows 2l and 2l+1 at kx = kb*B + c
    {
        constexpr int LPT = MAXE / 2;
        const int per_blk = np * B;
        const int nitems = g.nblk * per_blk;
        const float inv_pb = 1.0f / (float)per_blk, inv_b = 1.0f / (float)B;
        for (int base = 0; base < nitems; base += NT * LPT) {
            float2 ya[LPT], yb[LPT];
#pragma unroll
            for (int t = 0; t < LPT; ++t) {
                const int it = base + threadIdx.x + t * NT;
                ya[t] = make_float2(0.f, 0.f);
                yb[t] = ya[t];
                if (it < nitems) {
                    const int kb = fdiv(it, per_blk, inv_pb);
                    const int rem = it - kb * per_blk;
                    const int l = fdiv(rem, B, inv_b), c = rem - l * B;
                    const int k = kb * B + c;
                    if (k < H) {
                        const int64_t bse = ((int64_t)kb * N + a0 + 2 * l) * B + c;
                        ya[t] = u[bse];
                        if (2 * l + 1 < nr) yb[t] = u[bse + B];
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < LPT; ++t) {
                const int it = base + threadIdx.x + t * NT;
                if (it < nitems) {
                    const int kb = fdiv(it, per_blk, inv_pb);
                    const int rem = it - kb * per_blk;
                    const int l = fdiv(rem, B, inv_b), c = rem - l * B;
                    const int k = kb * B + c;
                    if (k < H) {
                        float2 a = ya[t], b = yb[t];
                        if (k == 0 || ((M & 1) == 0 && k == half)) { a.y = 0.f; b.y = 0.f; }
                        buf[lds_row(l, k, M)] = make_float2(a.x - b.y, a.y + b.x);
                        if (k >= 1 && k <= M - H) buf[lds_row(l, M - k, M)] = make_float2(a.x + b.y, b.x - a.y);
                    }
                }
            }
        }
    }
    __syncthreads();
    fft_lds<float, 1, NT, MAXE, false>(buf, ppb, M, plan, tw);
    float *out = dst + s.xoff + (int64_t)a0 * N;
    const float *pp = pvec + s.xoff + (int64_t)a0 * N;
    double acc = 0.0;
    if ((N & 1) == 0) {
        const int hn = N >> 1;
        const float inv = 1.0f / (float)hn;
        for (int it = threadIdx.x; it < nr * hn; it += NT) {
            const int row = fdiv(it, hn, inv), q = it - row * hn;
            const int n = 2 * q, l = row >> 1;
            const float2 z0 = buf[lds_row(l, n, M)], z1 = buf[lds_row(l, n + 1, M)];
            const float2 v = (row & 1) ? make_float2(z0.y, z1.y) : make_float2(z0.x, z1.x);
            *reinterpret_cast<float2 *>(out + (int64_t)row * N + n) = v;
            if (DOT) {
                const float2 pv = *reinterpret_cast<const float2 *>(pp + (int64_t)row * N + n);
                acc += (double)pv.x * (double)v.x + (double)pv.y * (double)v.y;
            }
        }
    } else {
        for (int it = threadIdx.x; it < nr * N; it += NT) {
            const int row = it / N, n = it - row * N;
            const float2 z = buf[lds_row(row >> 1, n, M)];
            const float v = (row & 1) ? z.y : z.x;
            out[(int64_t)row * N + n] = v;
            if (DOT) acc += (double)pp[(int64_t)row * N + n] * (double)v;
        }
    }
    if (DOT) {
        // block reduction in a fixed order
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        __shared__ double wsum[NT / 64];
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < NT / 64; ++w) t += wsum[w];
            partial[slot_base + blockIdx.y * gridDim.x + blockIdx.x] = t;
        }
    }
}

// ----------------------------------------------------------------------------
// launch configuration: a few compiled block shapes per pass; a variant picks
// one per size class (defaults below, or the autotuner's choice)
// ----------------------------------------------------------------------------
namespace {
struct RowShape { int nt, maxe; };
struct ColShape { int nt, maxe; bool pf; };
constexpr RowShape ROW_SHAPES[] = {{256, 8}, {256, 16}, {512, 8}};
constexpr ColShape COL_SHAPES[] = {{512, 16, false}, {512, 16, true}, {256, 16, false}, {256, 8, false}};
constexpr int N_ROW_SHAPES = 3, N_COL_SHAPES = 4;
} // namespace

int toep_num_row_shapes() { return N_ROW_SHAPES; }
int toep_num_col_shapes() { return N_COL_SHAPES; }
int toep_col_cap(int col) { return COL_SHAPES[col].nt * COL_SHAPES[col].maxe; }

ToepVariant toep_default_variant(int M) {
    ToepVariant v;
    v.row = 0;
    v.col = 0;
    v.B = std::max(1, std::min(toep_col_cap(0) / M, 16));
    return v;
}

void toep_configure(ToepClass &k, ToepVariant v) {
    LC_REQUIRE(v.row >= 0 && v.row < N_ROW_SHAPES && v.col >= 0 && v.col < N_COL_SHAPES,
               "toeplitz: bad kernel variant");
    const RowShape rs = ROW_SHAPES[v.row];
    const int npairs = (k.N + 1) / 2;
    const int cap_row = rs.nt * rs.maxe;
    LC_REQUIRE(k.M <= cap_row, "toeplitz: row shape too small for this size");
    k.ppb = std::max(1, std::min(cap_row / k.M, npairs));
    k.p1_blocks = ceil_div(npairs, k.ppb);
    const int cap_col = toep_col_cap(v.col);
    LC_REQUIRE(k.M <= cap_col, "toeplitz: column shape too small for this size");
    int B = v.B <= 0 ? std::max(1, std::min(cap_col / k.M, 16)) : v.B;
    B = std::max(1, std::min(B, std::min(cap_col / k.M, 16)));
    if ((k.N & 1) && (B & 1)) B = (B + 1 <= std::min(cap_col / k.M, 16)) ? B + 1 : B - 1;
    LC_REQUIRE(B >= 1 && !((k.N & 1) && (B & 1)), "toeplitz: no even block width fits");
    v.B = B;
    k.var = v;
    k.B = B;
    k.nblk = ceil_div(k.H, B);
    k.p2_blocks = k.nblk;
    k.g.N = k.N;
    k.g.M = k.M;
    k.g.H = k.H;
    k.g.B = B;
    k.g.nblk = k.nblk;
    k.g.ppb = k.ppb;
}

// U elements per slice, enough for any block width <= 16
int64_t toep_u_size(const ToepClass &k) { return ((int64_t)(k.H + 16) * k.N + 1) & ~(int64_t)1; }

#define LC_ROW_DISPATCH(KERNEL, FLAG, ...)                                                 \
    switch (k.var.row) {                                                                  \
    case 0: hipLaunchKernelGGL((KERNEL<256, 8, FLAG>), grid, dim3(256), sh, c.stream(), __VA_ARGS__); break; \
    case 1: hipLaunchKernelGGL((KERNEL<256, 16, FLAG>), grid, dim3(256), sh, c.stream(), __VA_ARGS__); break; \
    default: hipLaunchKernelGGL((KERNEL<512, 8, FLAG>), grid, dim3(512), sh, c.stream(), __VA_ARGS__); break; \
    }

void toep_launch_fwd(Ctx &c, const ToepClass &k, const CgState *st, const float *src,
                     const float *rsrc, float *pdst, float2 *U, bool fuse) {
    dim3 grid(k.p1_blocks, k.nslices);
    size_t sh = sizeof(float2) * lds_round(k.ppb * k.M);
    if (fuse) {
        LC_ROW_DISPATCH(k_toep_rows_fwd, true, k.slices.as<SliceRef>(), st, src, rsrc, pdst, U,
                        k.g, k.plan, k.tw);
    } else {
        LC_ROW_DISPATCH(k_toep_rows_fwd, false, k.slices.as<SliceRef>(), st, src, rsrc, pdst, U,
                        k.g, k.plan, k.tw);
    }
    LC_HIP(hipGetLastError());
}

void toep_launch_cols(Ctx &c, const ToepClass &k, const ProbDev *pd, const CgState *st,
                      float2 *U) {
    dim3 grid(k.p2_blocks, k.nslices);
    size_t sh = sizeof(float2) * lds_round(k.B * k.M);
    const SliceRef *sl = k.slices.as<SliceRef>();
    switch (k.var.col) {
    case 0: hipLaunchKernelGGL((k_toep_cols<512, 16, false>), grid, dim3(512), sh, c.stream(), sl, pd, st, U, k.g, k.plan, k.tw); break;
    case 1: hipLaunchKernelGGL((k_toep_cols<512, 16, true>), grid, dim3(512), sh, c.stream(), sl, pd, st, U, k.g, k.plan, k.tw); break;
    case 2: hipLaunchKernelGGL((k_toep_cols<256, 16, false>), grid, dim3(256), sh, c.stream(), sl, pd, st, U, k.g, k.plan, k.tw); break;
    default: hipLaunchKernelGGL((k_toep_cols<256, 8, false>), grid, dim3(256), sh, c.stream(), sl, pd, st, U, k.g, k.plan, k.tw); break;
    }
    LC_HIP(hipGetLastError());
}

void toep_launch_inv(Ctx &c, const ToepClass &k, const CgState *st, const float2 *U, float *dst,
                     const float *pvec, double *partial, bool dot) {
    dim3 grid(k.p1_blocks, k.nslices);
    size_t sh = sizeof(float2) * lds_round(k.ppb * k.M);
    if (dot) {
        LC_ROW_DISPATCH(k_toep_rows_inv, true, k.slices.as<SliceRef>(), st, U, dst, pvec, partial,
                        k.slot_base, k.g, k.plan, k.tw);
    } else {
        LC_ROW_DISPATCH(k_toep_rows_inv, false, k.slices.as<SliceRef>(), st, U, dst, pvec,
                        partial, k.slot_base, k.g, k.plan, k.tw);
    }
    LC_HIP(hipGetLastError());
}

} // namespace lc

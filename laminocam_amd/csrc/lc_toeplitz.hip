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
//               into two half spectra -> U (N x H complex, row major)
//   P2 columns: per column kx < H: load N values, FFT_M, x khat[kx][:],
//               IFFT_M, keep N values -> U in place
//   P3 rows:    pairs of half spectra rebuilt into one complex row by
//               Hermitian symmetry, IFFT_M, real/imag -> the two output rows
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

// P in periodic (centred) layout: index j <-> m = j (j < M/2) or j - M.
// Output complex double M x M with zero imaginary part.
__global__ void k_psf2d_gen(double2 *P, const double *theta, int nang, int nrad, int M) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    if (ix >= M) return;
    const double kx = (double)(ix < M / 2 ? ix : ix - M);
    const double ky = (double)(iy < M / 2 ? iy : iy - M);
    const double inv_n = 1.0 / (double)nrad;
    double s = 0.0;
    for (int i = 0; i < nang; ++i) {
        double sn, cs;
        sincos(theta[i], &sn, &cs);
        s += dirichlet_d(kx * cs + ky * sn, nrad, inv_n);
    }
    P[(size_t)iy * M + ix] = make_double2(s, 0.0);
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
        const int l = idx / L;
        buf[idx] = (l < nloc) ? data[(size_t)(row0 + l) * L + (idx - l * L)] : cpx<T>{};
    }
    __syncthreads();
    fft_lds<T, SIGN, NT, MAXE, false>(buf, nl, L, 1, plan, tw);
    for (int idx = threadIdx.x; idx < nloc * L; idx += NT) {
        const int l = idx / L;
        data[(size_t)(row0 + l) * L + (idx - l * L)] = buf[idx];
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
        buf[idx] = (col0 + c < H) ? P[(size_t)i * M + col0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, -1, NT, MAXE, true>(buf, B, 1, B, plan, tw);
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int c = idx / M, i = idx - c * M;
        if (col0 + c < H) khatT[(size_t)(col0 + c) * M + i] = (float)(buf[i * B + c].x * scale);
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
    g.psf_ready = false;
}

void geom2d_build_psf(Ctx &c, Geom2D &g) {
    const int M = (int)g.M, H = (int)g.H;
    FftPlan plan;
    LC_REQUIRE(fft_make_plan(M, plan), "PSF size is not 2^p 3^q");
    double2 *P = reinterpret_cast<double2 *>(c.scratch(0, sizeof(double2) * (size_t)M * M));
    hipLaunchKernelGGL(k_psf2d_gen, dim3(ceil_div(M, 256), M), dim3(256), 0, c.stream(), P,
                       g.d_theta.as<double>(), (int)g.nang, (int)g.nrad, M);
    LC_HIP(hipGetLastError());
    const double2 *tw = c.twiddles<double>(M);
    constexpr int NT = 256, MAXE = 16;
    // rows: nl lines per block with nl*M <= NT*MAXE
    const int nl = std::max(1, NT * MAXE / M);
    hipLaunchKernelGGL((k_rows_c2c<double, -1, NT, MAXE>), dim3(ceil_div(M, nl)), dim3(NT),
                       sizeof(double2) * nl * M, c.stream(), P, M, nl, plan, tw);
    LC_HIP(hipGetLastError());
    g.khatT.alloc(sizeof(float) * (size_t)H * M);
    const int B = std::max(1, NT * MAXE / M);
    const double scale = 1.0 / ((double)M * (double)M * (double)g.N);
    hipLaunchKernelGGL((k_psf_cols<NT, MAXE>), dim3(ceil_div(H, B)), dim3(NT),
                       sizeof(double2) * B * M, c.stream(), P, g.khatT.as<float>(), M, H, B, plan,
                       tw, scale);
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

// P1: U rows of one slice from pairs of real rows.  FUSE: the row source is
// the CG direction p = r + beta p (written back) instead of a plain input.
template <int NT, int MAXE, bool FUSE>
__global__ void __launch_bounds__(NT)
k_toep_rows_fwd(const SliceRef *__restrict__ sl, const CgState *__restrict__ st,
                const float *src, const float *__restrict__ rsrc, float *pdst,
                float2 *__restrict__ U, int N, int M, int H, int ppb, FftPlan plan,
                const float2 *__restrict__ tw) {
    const SliceRef s = sl[blockIdx.y];
    if (!prob_active(st, s.prob)) return;
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    const int pair0 = blockIdx.x * ppb;
    const float *x = src + s.xoff;
    float beta = 0.0f;
    if (FUSE) beta = (float)st[s.prob].beta;
    for (int idx = threadIdx.x; idx < ppb * M; idx += NT) {
        const int l = idx / M, n = idx - l * M;
        const int a = 2 * (pair0 + l);
        float re = 0.0f, im = 0.0f;
        if (n < N) {
            if (a < N) {
                const int64_t q = (int64_t)a * N + n;
                if (FUSE) {
                    re = rsrc[s.xoff + q] + beta * x[q];
                    pdst[s.xoff + q] = re;
                } else {
                    re = x[q];
                }
            }
            if (a + 1 < N) {
                const int64_t q = (int64_t)(a + 1) * N + n;
                if (FUSE) {
                    im = rsrc[s.xoff + q] + beta * x[q];
                    pdst[s.xoff + q] = im;
                } else {
                    im = x[q];
                }
            }
        }
        buf[idx] = make_float2(re, im);
    }
    __syncthreads();
    fft_lds<float, -1, NT, MAXE, false>(buf, ppb, M, 1, plan, tw);
    float2 *u = U + s.uoff;
    for (int idx = threadIdx.x; idx < ppb * H; idx += NT) {
        const int l = idx / H, k = idx - l * H;
        const int a = 2 * (pair0 + l);
        if (a >= N) continue;
        const float2 z = buf[l * M + k];
        const float2 zc = cconj(buf[l * M + (k == 0 ? 0 : M - k)]);
        // X_a = (Z + conj Z(-k))/2,  X_b = (Z - conj Z(-k))/(2i)
        u[(int64_t)a * H + k] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
        if (a + 1 < N) u[(int64_t)(a + 1) * H + k] = make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
    }
}

// P2: column FFT, multiply by the real spectrum, inverse FFT, crop to N rows
template <int NT, int MAXE>
__global__ void __launch_bounds__(NT)
k_toep_cols(const SliceRef *__restrict__ sl, const ProbDev *__restrict__ pd,
            const CgState *__restrict__ st, float2 *__restrict__ U, int N, int M, int H, int B,
            int nbx, FftPlan plan, const float2 *__restrict__ tw) {
    const int lin = blockIdx.x + blockIdx.y * gridDim.x;
    const int logical = xcd_remap(lin, gridDim.x * gridDim.y);
    const int sidx = logical / nbx, bx = logical - sidx * nbx;
    const SliceRef s = sl[sidx];
    if (!prob_active(st, s.prob)) return;
    const float *__restrict__ kh = pd[s.prob].khatT;
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    float2 *u = U + s.uoff;
    const int col0 = bx * B;
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        buf[idx] = (i < N && col0 + c < H) ? u[(int64_t)i * H + col0 + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<float, -1, NT, MAXE, true>(buf, B, 1, B, plan, tw);
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int c = idx / M, i = idx - c * M;
        const float k = (col0 + c < H) ? kh[(int64_t)(col0 + c) * M + i] : 0.0f;
        buf[i * B + c] = cscale(buf[i * B + c], k);
    }
    __syncthreads();
    fft_lds<float, 1, NT, MAXE, true>(buf, B, 1, B, plan, tw);
    for (int idx = threadIdx.x; idx < N * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        if (col0 + c < H) u[(int64_t)i * H + col0 + c] = buf[idx];
    }
}

// P3: rebuild full spectra of two rows, inverse FFT, write real rows.
// DOT: also the block's partial of <p, Ap> into partial[slot].
template <int NT, int MAXE, bool DOT>
__global__ void __launch_bounds__(NT)
k_toep_rows_inv(const SliceRef *__restrict__ sl, const CgState *__restrict__ st,
                const float2 *__restrict__ U, float *__restrict__ dst,
                const float *__restrict__ pvec, double *__restrict__ partial, int slot_base,
                int N, int M, int H, int ppb, FftPlan plan, const float2 *__restrict__ tw) {
    const SliceRef s = sl[blockIdx.y];
    if (!prob_active(st, s.prob)) return;
    float2 *buf = reinterpret_cast<float2 *>(lc_smem);
    const int pair0 = blockIdx.x * ppb;
    const float2 *u = U + s.uoff;
    const int half = M / 2;
    for (int idx = threadIdx.x; idx < ppb * M; idx += NT) {
        const int l = idx / M, k = idx - l * M;
        const int a = 2 * (pair0 + l);
        float2 ya = make_float2(0.f, 0.f), yb = make_float2(0.f, 0.f);
        if (a < N) {
            const bool lo = k < H;
            const int kk = lo ? k : M - k;
            ya = u[(int64_t)a * H + kk];
            if (a + 1 < N) yb = u[(int64_t)(a + 1) * H + kk];
            if (kk == 0 || ((M % 2 == 0) && kk == half)) { ya.y = 0.f; yb.y = 0.f; }
            if (!lo) { ya.y = -ya.y; yb.y = -yb.y; }
        }
        buf[idx] = make_float2(ya.x - yb.y, ya.y + yb.x);
    }
    __syncthreads();
    fft_lds<float, 1, NT, MAXE, false>(buf, ppb, M, 1, plan, tw);
    float *out = dst + s.xoff;
    double acc = 0.0;
    for (int idx = threadIdx.x; idx < ppb * N; idx += NT) {
        const int l = idx / N, n = idx - l * N;
        const int a = 2 * (pair0 + l);
        if (a >= N) continue;
        const float2 z = buf[l * M + n];
        const int64_t q0 = (int64_t)a * N + n;
        out[q0] = z.x;
        if (DOT) acc += (double)pvec[s.xoff + q0] * (double)z.x;
        if (a + 1 < N) {
            const int64_t q1 = q0 + N;
            out[q1] = z.y;
            if (DOT) acc += (double)pvec[s.xoff + q1] * (double)z.y;
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

// explicit launch helpers used by the engine
struct ToepLaunch {
    static constexpr int NT = 256;
    static constexpr int MAXE_ROW = 16;
    static constexpr int MAXE_COL = 32;
};

void toep_configure(ToepClass &k) {
    const int cap_row = ToepLaunch::NT * ToepLaunch::MAXE_ROW;
    const int cap_col = ToepLaunch::NT * ToepLaunch::MAXE_COL;
    const int npairs = (k.N + 1) / 2;
    k.ppb = std::max(1, std::min(cap_row / k.M, npairs));
    k.p1_blocks = ceil_div(npairs, k.ppb);
    k.B = std::max(1, std::min(cap_col / k.M, 16));
    k.p2_blocks = ceil_div(k.H, k.B);
}

void toep_launch_fwd(Ctx &c, const ToepClass &k, const CgState *st, const float *src,
                     const float *rsrc, float *pdst, float2 *U, bool fuse) {
    constexpr int NT = ToepLaunch::NT, ME = ToepLaunch::MAXE_ROW;
    dim3 grid(k.p1_blocks, k.nslices);
    size_t sh = sizeof(float2) * (size_t)k.ppb * k.M;
    if (fuse)
        hipLaunchKernelGGL((k_toep_rows_fwd<NT, ME, true>), grid, dim3(NT), sh, c.stream(),
                           k.slices.as<SliceRef>(), st, src, rsrc, pdst, U, k.N, k.M, k.H, k.ppb,
                           k.plan, k.tw);
    else
        hipLaunchKernelGGL((k_toep_rows_fwd<NT, ME, false>), grid, dim3(NT), sh, c.stream(),
                           k.slices.as<SliceRef>(), st, src, rsrc, pdst, U, k.N, k.M, k.H, k.ppb,
                           k.plan, k.tw);
    LC_HIP(hipGetLastError());
}

void toep_launch_cols(Ctx &c, const ToepClass &k, const ProbDev *pd, const CgState *st,
                      float2 *U) {
    constexpr int NT = ToepLaunch::NT, ME = ToepLaunch::MAXE_COL;
    dim3 grid(k.p2_blocks, k.nslices);
    size_t sh = sizeof(float2) * (size_t)k.B * k.M;
    hipLaunchKernelGGL((k_toep_cols<NT, ME>), grid, dim3(NT), sh, c.stream(),
                       k.slices.as<SliceRef>(), pd, st, U, k.N, k.M, k.H, k.B, k.p2_blocks, k.plan,
                       k.tw);
    LC_HIP(hipGetLastError());
}

void toep_launch_inv(Ctx &c, const ToepClass &k, const CgState *st, const float2 *U, float *dst,
                     const float *pvec, double *partial, bool dot) {
    constexpr int NT = ToepLaunch::NT, ME = ToepLaunch::MAXE_ROW;
    dim3 grid(k.p1_blocks, k.nslices);
    size_t sh = sizeof(float2) * (size_t)k.ppb * k.M;
    if (dot)
        hipLaunchKernelGGL((k_toep_rows_inv<NT, ME, true>), grid, dim3(NT), sh, c.stream(),
                           k.slices.as<SliceRef>(), st, U, dst, pvec, partial, k.slot_base, k.N,
                           k.M, k.H, k.ppb, k.plan, k.tw);
    else
        hipLaunchKernelGGL((k_toep_rows_inv<NT, ME, false>), grid, dim3(NT), sh, c.stream(),
                           k.slices.as<SliceRef>(), st, U, dst, pvec, partial, k.slot_base, k.N,
                           k.M, k.H, k.ppb, k.plan, k.tw);
    LC_HIP(hipGetLastError());
}

} // namespace lc

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

// Any circular-convolution size >= 2n - 1 gives the same linear convolution
// after the crop: the PSF values at |m| <= n - 1 do not depend on it, and the
// unnormalised FFT round trip's factor is divided out with it.  The FFT
// lengths compiled into the Toeplitz kernels run much faster than runtime
// plans (1458 = 2 3^6 takes seven radix stages, 1536 four), so take the
// smallest compiled length within 25% of next_fast_dim(2n - 1), else that.
int64_t toep_conv_size(int64_t n) {
    const int64_t m = next_fast_dim(2 * n - 1);
    static const int64_t compiled[] = {256, 384, 512, 768, 1024, 1536, 2048};
    for (int64_t L : compiled)
        if (L >= m) return L * 4 <= m * 5 ? L : m;
    return m;
}

// ----------------------------------------------------------------------------
// FP64 PSF: P(m) = sum_i D_nrad(m_x cos th_i + m_y sin th_i),
// D_n(t) = sin(pi t)/sin(pi t/n) (the radial sum of exp(i r_k t) over the
// polar grid's r_k = (k+1/2) 2pi/n - pi); dirichlet_d evaluates it without
// dividing 0 by 0.
// ----------------------------------------------------------------------------
constexpr int PSF_RUN = 8;          // consecutive kx per thread (rotation recurrence)
constexpr int PSF_ANG_CHUNK = 512;  // angles staged in LDS per pass

// P in periodic (centred) layout: index j <-> m = j (j < M/2) or j - M.
// Rows iy <= M/2 are computed and mirrored through P(-m) = P(m).  Along a run
// of consecutive kx the phases exp(i pi t) and exp(i pi t / n) advance by a
// fixed rotation per step; the ratio of their imaginary parts is D_n(t).
// Where sin(pi t / n) is small the direct reduced form is used instead.
__global__ void __launch_bounds__(256)
k_psf2d_gen(double2 *P, const double *theta, int nang, int nrad, int M, int full) {
    __shared__ double4 ang[PSF_ANG_CHUNK];    // (cos, sin, cos pi c, sin pi c)
    __shared__ double2 ang2[PSF_ANG_CHUNK];   // (cos pi c/n, sin pi c/n)
    const double inv_n = 1.0 / (double)nrad;
    const int runs = (M + PSF_RUN - 1) / PSF_RUN;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = gid / runs;
    const int run = gid - iy * runs;
    // full: every row computed (no mirror), as the exported kernel_hat needs
    const bool valid = full ? iy < M : iy <= M / 2;   // no early return: the block synchronises below
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
    const bool mirror = !full && iy >= 1 && 2 * iy < M;
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
// khatT[kx][ky] = Re(FFT)[ky][kx] * scale (stored as OT: float or double)
template <typename OT, int NT, int MAXE>
__global__ void __launch_bounds__(NT) k_psf_cols(const double2 *P, OT *khatT, int M, int H, int B,
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
        if (col0 + c < H) khatT[(size_t)(col0 + c) * M + i] = (OT)(buf[lds_col(c, i, B)].x * scale);
    }
}

// reference-layout kernel_hat: FP64 column FFT of the row-transformed full
// PSF (centred layout) for kx < H, times the phase of the reference's storage
// with the centre at index h = floor(M/2): exp(-2 pi i (kx + ky) h / M).
// For even M the row and column m = -M/2 have no +M/2 partner, so this
// spectrum is complex; those entries never reach the cropped convolution.
template <typename OT, int NT, int MAXE>
__global__ void __launch_bounds__(NT) k_khat_export(const double2 *P, cpx<OT> *out, int M, int H,
                                                    int B, FftPlan plan,
                                                    const double2 *__restrict__ tw) {
    double2 *buf = reinterpret_cast<double2 *>(lc_smem);
    const int col0 = blockIdx.x * B;
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        buf[swz(idx)] = (col0 + c < H) ? P[(size_t)i * M + col0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, -1, NT, MAXE, true>(buf, B, 0, plan, tw);
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int ky = idx / B, c = idx - ky * B;
        const int kx = col0 + c;
        if (kx >= H) continue;
        const double2 v = buf[lds_col(c, ky, B)];
        const long long ph = ((long long)(kx + ky) * (M / 2)) % M;
        double s, co;
        sincospi(-2.0 * (double)ph / (double)M, &s, &co);
        out[(size_t)ky * H + kx] = cmake<cpx<OT>>((OT)(v.x * co - v.y * s), (OT)(v.x * s + v.y * co));
    }
}

void geom2d_init(Ctx &c, Geom2D &g, const double *theta, int64_t nang, int64_t nrad,
                 int64_t image_n) {
    LC_REQUIRE(nang >= 1 && nrad >= 2 && image_n >= 2, "geometry: need nang>=1, nrad>=2, N>=2");
    LC_REQUIRE(nang <= (1 << 20), "geometry: too many angles");
    g.nang = nang;
    g.nrad = nrad;
    g.N = image_n;
    g.Mr = next_fast_dim(2 * image_n - 1);
    g.M = toep_conv_size(image_n);
    g.H = g.M / 2 + 1;
    LC_REQUIRE(g.M <= 4096, "geometry: image size above 2048 is not supported");
    g.theta.assign(theta, theta + nang);
    g.d_theta.alloc(sizeof(double) * nang);
    LC_HIP(hipMemcpyAsync(g.d_theta.as(), theta, sizeof(double) * nang, hipMemcpyHostToDevice,
                          c.stream()));
    c.sync();
    g.psf_ready = false;
    g.psf64_ready = false;
}

void geom2d_build_psf(Ctx &c, Geom2D &g, bool f64) {
    const int M = (int)g.M, H = (int)g.H;
    FftPlan plan;
    LC_REQUIRE(fft_make_plan(M, plan), "PSF size is not 2^p 3^q");
    double2 *P = reinterpret_cast<double2 *>(c.scratch(0, sizeof(double2) * (size_t)M * M));
    const int runs = (M + PSF_RUN - 1) / PSF_RUN;
    const int threads = (M / 2 + 1) * runs;
    hipLaunchKernelGGL(k_psf2d_gen, dim3(ceil_div(threads, 256)), dim3(256), 0, c.stream(), P,
                       g.d_theta.as<double>(), (int)g.nang, (int)g.nrad, M, 0);
    LC_HIP(hipGetLastError());
    const double2 *tw = c.twiddles<double>(M);
    constexpr int NT = 256, MAXE = 16;
    // rows: nl lines per block with nl*M <= NT*MAXE
    const int nl = std::max(1, NT * MAXE / M);
    hipLaunchKernelGGL((k_rows_c2c<double, -1, NT, MAXE>), dim3(ceil_div(M, nl)), dim3(NT),
                       sizeof(double2) * lds_round(nl * M), c.stream(), P, M, nl, plan, tw);
    LC_HIP(hipGetLastError());
    const int B = std::max(1, NT * MAXE / M);
    const double scale = 1.0 / ((double)M * (double)M * (double)g.N);
    if (f64) {
        g.khatT64.alloc(sizeof(double) * (size_t)H * M);
        hipLaunchKernelGGL((k_psf_cols<double, NT, MAXE>), dim3(ceil_div(H, B)), dim3(NT),
                           sizeof(double2) * lds_round(B * M), c.stream(), P,
                           g.khatT64.as<double>(), M, H, B, plan, tw, scale);
    } else {
        g.khatT.alloc(sizeof(float) * (size_t)H * M);
        hipLaunchKernelGGL((k_psf_cols<float, NT, MAXE>), dim3(ceil_div(H, B)), dim3(NT),
                           sizeof(double2) * lds_round(B * M), c.stream(), P, g.khatT.as<float>(),
                           M, H, B, plan, tw, scale);
    }
    LC_HIP(hipGetLastError());
    // the apply's FFT twiddles, so that a solve on this context finds them
    // without allocating
    if (f64) (void)c.twiddles<double>((int)g.M);
    else (void)c.twiddles<float>((int)g.M);
    // P lives in a context scratch slot that the next call may reuse
    c.sync();
    if (f64) g.psf64_ready = true;
    else g.psf_ready = true;
}

template <typename OT> void geom2d_kernel_hat(Ctx &c, Geom2D &g, OT *out_host) {
    // the reference's layout and size, independent of the convolution size
    const int M = (int)g.Mr, H = M / 2 + 1;
    FftPlan plan;
    LC_REQUIRE(fft_make_plan(M, plan), "PSF size is not 2^p 3^q");
    double2 *P = reinterpret_cast<double2 *>(c.scratch(0, sizeof(double2) * (size_t)M * M));
    const int runs = (M + PSF_RUN - 1) / PSF_RUN;
    hipLaunchKernelGGL(k_psf2d_gen, dim3(ceil_div((int64_t)M * runs, 256)), dim3(256), 0, c.stream(),
                       P, g.d_theta.as<double>(), (int)g.nang, (int)g.nrad, M, 1);
    LC_HIP(hipGetLastError());
    const double2 *tw = c.twiddles<double>(M);
    constexpr int NT = 256, MAXE = 16;
    const int nl = std::max(1, NT * MAXE / M);
    hipLaunchKernelGGL((k_rows_c2c<double, -1, NT, MAXE>), dim3(ceil_div(M, nl)), dim3(NT),
                       sizeof(double2) * lds_round(nl * M), c.stream(), P, M, nl, plan, tw);
    LC_HIP(hipGetLastError());
    cpx<OT> *tmp = reinterpret_cast<cpx<OT> *>(c.scratch(1, sizeof(cpx<OT>) * (size_t)M * H));
    const int B = std::max(1, NT * MAXE / M);
    hipLaunchKernelGGL((k_khat_export<OT, NT, MAXE>), dim3(ceil_div(H, B)), dim3(NT),
                       sizeof(double2) * lds_round(B * M), c.stream(), P, tmp, M, H, B, plan, tw);
    LC_HIP(hipGetLastError());
    LC_HIP(hipMemcpyAsync(out_host, tmp, sizeof(cpx<OT>) * (size_t)M * H, hipMemcpyDeviceToHost,
                          c.stream()));
    c.sync();
}
template void geom2d_kernel_hat<float>(Ctx &, Geom2D &, float *);
template void geom2d_kernel_hat<double>(Ctx &, Geom2D &, double *);

// ----------------------------------------------------------------------------
// launch configuration: a few compiled block shapes per pass; a variant picks
// one per size class (defaults below, or the autotuner's choice).  Lengths
// with a compile-time instantiation use it; others take the runtime plan.
// ----------------------------------------------------------------------------
namespace {
struct RowShape { int nt, maxe; };
struct ColShape { int nt, maxe; };
constexpr RowShape ROW_SHAPES[] = {{256, 8}, {512, 8}, {256, 16}};
constexpr ColShape COL_SHAPES[] = {{256, 16}, {256, 8}, {512, 16}};
constexpr int N_ROW_SHAPES = 3, N_COL_SHAPES = 3;
} // namespace

#define LC_DECL_L(L)                                                                              \
    template <typename R>                                                                         \
    void toep_fwd_L##L(Ctx &, const ToepClass &, const CgState *, const R *, const R *, R *,      \
                       cpx<R> *, bool);                                                           \
    template <typename R>                                                                         \
    void toep_cols_L##L(Ctx &, const ToepClass &, const ProbDev *, const CgState *, cpx<R> *);    \
    template <typename R>                                                                         \
    void toep_inv_L##L(Ctx &, const ToepClass &, const CgState *, const cpx<R> *, R *, const R *, \
                       double *, bool);
LC_DECL_L(0)
LC_DECL_L(256)
LC_DECL_L(384)
LC_DECL_L(512)
LC_DECL_L(768)
LC_DECL_L(1024)
LC_DECL_L(1536)
LC_DECL_L(2048)

int toep_num_row_shapes() { return N_ROW_SHAPES; }
int toep_num_col_shapes() { return N_COL_SHAPES; }
int toep_col_cap(int col) { return COL_SHAPES[col].nt * COL_SHAPES[col].maxe; }
// 16-byte complex: shapes whose LDS block stays within 64 KiB (shape 2 does not)
int toep_num_col_shapes_for(bool f64) { return f64 ? 2 : N_COL_SHAPES; }

static int pow2_floor(int v) {
    int p = 1;
    while (2 * p <= v) p *= 2;
    return p;
}

ToepVariant toep_default_variant(int M) {
    ToepVariant v;
    v.row = 0;
    v.col = 1;
    v.B = 0;
    if (M > toep_col_cap(1)) v.col = 0;
    if (M > ROW_SHAPES[0].nt * ROW_SHAPES[0].maxe) v.row = 1;
    return v;
}

ToepVariant toep_fit_variant(ToepVariant v, int N, int M, bool f64) {
    if (!(N & 1) || v.col < 0 || v.col >= N_COL_SHAPES || toep_col_cap(v.col) >= 2 * M) return v;
    for (int col : {0, 2, 1}) {
        if (col < toep_num_col_shapes_for(f64) && toep_col_cap(col) >= 2 * M) {
            v.col = col;
            v.B = 0;
            return v;
        }
    }
    return v;   // toep_configure reports the size as unsupported
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
    // columns per block: a power of two (interleaved LDS lines), even when N
    // is odd so that N*B stays even for the float4 block copies
    const int bmax = pow2_floor(std::min(cap_col / k.M, 16));
    int B = v.B <= 0 ? bmax : pow2_floor(std::min(v.B, bmax));
    if ((k.N & 1) && B < 2) B = 2;
    LC_REQUIRE(B <= cap_col / k.M, "toeplitz: no block width fits this size");
    int lsh = 0;
    while ((1 << lsh) < B) ++lsh;
    v.B = B;
    k.var = v;
    k.B = B;
    k.nblk = ceil_div(k.H, B);
    k.p2_blocks = k.nblk;
    k.g.N = k.N;
    k.g.M = k.M;
    k.g.H = k.H;
    k.g.B = B;
    k.g.lsh = lsh;
    k.g.nblk = k.nblk;
    k.g.ppb = k.ppb;
}

// 3-D classes: row shape by M (512 x 8 above 1024: 1.7% faster per
// laminography CG iteration at M = 1536 than 256 x 8); the column shape only
// bounds the block width (4096 FP64 or 8192 FP32 elements per block)
ToepVariant toep3d_variant(int M, bool f64) {
    ToepVariant v;
    v.row = M > 1024 ? 1 : 0;
    v.col = f64 ? 0 : 2;
    v.B = 0;
    return v;
}

int toep3d_block_width(int N, int M, bool f64) {
    ToepClass k;
    k.N = N;
    k.M = M;
    k.H = M / 2 + 1;
    toep_configure(k, toep3d_variant(M, f64));
    return k.B;
}

// U elements per slice, enough for any block width <= 16
int64_t toep_u_size(const ToepClass &k) { return ((int64_t)(k.H + 16) * k.N + 1) & ~(int64_t)1; }

#define LC_DISPATCH_M(CALL)                                                                       \
    switch (k.M) {                                                                                \
    case 256: CALL(256); break;                                                                   \
    case 384: CALL(384); break;                                                                   \
    case 512: CALL(512); break;                                                                   \
    case 768: CALL(768); break;                                                                   \
    case 1024: CALL(1024); break;                                                                 \
    case 1536: CALL(1536); break;                                                                 \
    case 2048: CALL(2048); break;                                                                 \
    default: CALL(0); break;                                                                      \
    }

template <typename R>
void toep_launch_fwd(Ctx &c, const ToepClass &k, const CgState *st, const R *src, const R *rsrc,
                     R *pdst, cpx<R> *U, bool fuse) {
#define LC_F(L) toep_fwd_L##L<R>(c, k, st, src, rsrc, pdst, U, fuse)
    LC_DISPATCH_M(LC_F)
#undef LC_F
}

template <typename R>
void toep_launch_cols(Ctx &c, const ToepClass &k, const ProbDev *pd, const CgState *st, cpx<R> *U) {
#define LC_F(L) toep_cols_L##L<R>(c, k, pd, st, U)
    LC_DISPATCH_M(LC_F)
#undef LC_F
}

template <typename R>
void toep_launch_inv(Ctx &c, const ToepClass &k, const CgState *st, const cpx<R> *U, R *dst,
                     const R *pvec, double *partial, bool dot) {
#define LC_F(L) toep_inv_L##L<R>(c, k, st, U, dst, pvec, partial, dot)
    LC_DISPATCH_M(LC_F)
#undef LC_F
}

#define LC_LAUNCH_INST(R)                                                                         \
    template void toep_launch_fwd<R>(Ctx &, const ToepClass &, const CgState *, const R *,        \
                                     const R *, R *, cpx<R> *, bool);                             \
    template void toep_launch_cols<R>(Ctx &, const ToepClass &, const ProbDev *, const CgState *, \
                                      cpx<R> *);                                                  \
    template void toep_launch_inv<R>(Ctx &, const ToepClass &, const CgState *, const cpx<R> *,   \
                                     R *, const R *, double *, bool);
LC_LAUNCH_INST(float)
LC_LAUNCH_INST(double)
#undef LC_LAUNCH_INST

} // namespace lc

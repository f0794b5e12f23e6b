// 3-D laminography geometry, FP64 PSF and the middle pass of the 3-D
// Toeplitz normal operator on gfx950.
//
// Reference semantics (include/toeplitz.h:61-129): for a {n1, n2, n3} volume
// the PSF lives on dims_ = next_fast_dim(2 n - 1) per axis, P(m) = sum_j
// exp(+i m.w_j) over every grid point (no mask), and
// convolve(x) = crop_{dims_/2}(C2R(R2C(pad_right(x)) * khat)) / (prod(dims_) n2 n3).
// As in 2-D the PSF is real and even, so the spectrum of the centred PSF is
// real and the crop moves to offset 0; the scale is folded into the spectrum.
//
// With unit weights the PSF factorises per projection (the grid of one
// projection is a lattice qX_k u_i + qY_j v_i on a plane):
//   P(m) = sum_i D_ncols(m.u_i) D_nrows(m.v_i),  D_n(t) = sin(pi t)/sin(pi t/n).
//
// Apply on a {n1, N, N} volume (n2 = n3 = N, M = dims_[1] = dims_[2]):
//   P1  rows of every slice: the 2-D row pass (row pairs, FFT_M, half spectra) -> U
//   P2f every slice, every column block: FFT_M down the columns -> W (all M rows)
//   P2b every (k2, column block): FFT_M1 along the slices (zero padded),
//       x khat3 (k2, k3, :), inverse, keep n1 -> W
//   P2i every slice, column block: inverse FFT_M down the columns, keep N -> U
//   P3  rows of every slice: the 2-D inverse row pass -> output
// W is [nblk][n1][M][B] complex: the column passes read and write one
// contiguous M x B block per (column block, slice), and P2b reads, per slice,
// kb2 x B consecutive values (its lines) as one contiguous piece.  khat3 is
// blocked the same way ([nblk][M1][M][B] real), so P2b reads each slice
// index's spectrum values of its lines as one contiguous piece too.
#include <cmath>

#include <type_traits>

#include "lc_engine.hpp"

namespace lc {

extern __shared__ __align__(16) char lc_smem[];

// RotationTranspose(theta, gamma, beta) of include/rotation.h:28-56, columns 0 and 1
static void rotation_uv(double th, double ga, double be, double *uv) {
    const double ct = std::cos(th), st = std::sin(th);
    const double cb = std::cos(be), sb = std::sin(be);
    const double cg = std::cos(ga), sg = std::sin(ga);
    // column 0
    uv[0] = cb * cg;
    uv[1] = -cb * sg;
    uv[2] = sb;
    // column 1
    uv[3] = ct * sg + st * sb * cg;
    uv[4] = ct * cg - st * sb * sg;
    uv[5] = -st * cb;
}

void geom3d_init(Ctx &c, Geom3D &g, const double *theta, const double *gamma, const double *beta,
                 int64_t nproj, int64_t nrows, int64_t ncols, const int64_t *recon_dims) {
    LC_REQUIRE(nproj >= 1 && nrows >= 1 && ncols >= 1, "3-D geometry: empty grid");
    LC_REQUIRE(nproj <= (1 << 20), "3-D geometry: too many projections");
    LC_REQUIRE(recon_dims[0] >= 1 && recon_dims[1] >= 2 && recon_dims[2] >= 2,
               "3-D geometry: bad recon dims");
    g.nproj = nproj;
    g.nrows = nrows;
    g.ncols = ncols;
    g.n1 = recon_dims[0];
    g.n2 = recon_dims[1];
    g.n3 = recon_dims[2];
    g.N = recon_dims[1];
    g.M1 = next_fast_dim(2 * g.n1 - 1);
    g.Mr = next_fast_dim(2 * g.N - 1);
    g.M = toep_conv_size(g.N);
    g.H = g.M / 2 + 1;
    LC_REQUIRE(g.M <= 4096 && g.M1 <= 4096, "3-D geometry: volume too large");
    g.rot.resize(6 * nproj);
    for (int64_t i = 0; i < nproj; ++i) rotation_uv(theta[i], gamma[i], beta[i], &g.rot[6 * i]);
    g.d_rot.alloc(sizeof(double) * 6 * nproj);
    LC_HIP(hipMemcpyAsync(g.d_rot.as(), g.rot.data(), sizeof(double) * 6 * nproj,
                          hipMemcpyHostToDevice, c.stream()));
    c.sync();
    g.psf_ready = false;
    g.psf64_ready = false;
}

// ---------------------------------------------------------------------------
// FP64 PSF on the centred periodic index space (index a <-> m = a for
// a < (L+1)/2, a - L otherwise).  One thread per (iz, iy, run of P3_RUN ix) for
// iz <= M1/2; iz strictly inside (0, M1/2) also writes P(-m) = P(m).
// Along a run both Dirichlet arguments advance by fixed steps, so their
// numerator and denominator phases advance by fixed rotations.
// ---------------------------------------------------------------------------
constexpr int P3_RUN = 16;
constexpr int P3_CHUNK = 256;

__global__ void __launch_bounds__(256)
k_psf3d_gen(double2 *P, const double *rot, int nproj, int nrows, int ncols, int M1, int M,
            int full) {
    __shared__ double uvs[P3_CHUNK][6];
    __shared__ double4 stp[P3_CHUNK][2];   // (cos pi ux, sin pi ux, cos pi ux/nc, sin pi ux/nc), same for vx/nr
    const double inv_c = 1.0 / (double)ncols, inv_r = 1.0 / (double)nrows;
    const int runs = (M + P3_RUN - 1) / P3_RUN;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per_z = (int64_t)M * runs;
    const int iz = (int)(gid / per_z);
    const int rem = (int)(gid - (int64_t)iz * per_z);
    const int iy = rem / runs, run = rem - iy * runs;
    // full: every plane computed (no mirror), as the exported kernel_hat needs
    const bool valid = full ? iz < M1 : iz <= M1 / 2;   // no early return: the block synchronises below
    const int ix0 = run * P3_RUN;
    const double mz = (double)(iz < (M1 + 1) / 2 ? iz : iz - M1);
    const double my = (double)(iy < (M + 1) / 2 ? iy : iy - M);
    double acc[P3_RUN];
#pragma unroll
    for (int j = 0; j < P3_RUN; ++j) acc[j] = 0.0;
    for (int a0 = 0; a0 < nproj; a0 += P3_CHUNK) {
        const int na = min(P3_CHUNK, nproj - a0);
        __syncthreads();
        for (int i = threadIdx.x; i < na; i += blockDim.x) {
            const double *r = rot + 6 * (int64_t)(a0 + i);
            for (int q = 0; q < 6; ++q) uvs[i][q] = r[q];
            double s1, c1, s2, c2, s3, c3, s4, c4;
            sincospi(r[0], &s1, &c1);
            sincospi(r[0] * inv_c, &s2, &c2);
            sincospi(r[3], &s3, &c3);
            sincospi(r[3] * inv_r, &s4, &c4);
            stp[i][0] = make_double4(c1, s1, c2, s2);
            stp[i][1] = make_double4(c3, s3, c4, s4);
        }
        __syncthreads();
        if (!valid) continue;
        for (int i = 0; i < na; ++i) {
            const double ux = uvs[i][0], uy = uvs[i][1], uz = uvs[i][2];
            const double vx = uvs[i][3], vy = uvs[i][4], vz = uvs[i][5];
            const double4 su = stp[i][0], sv = stp[i][1];
            double a1r = 0, a1i = 0, b1r = 0, b1i = 0, a2r = 0, a2i = 0, b2r = 0, b2i = 0;
#pragma unroll
            for (int j = 0; j < P3_RUN; ++j) {
                const int ix = ix0 + j;
                if (ix >= M) break;
                const double mx = (double)(ix < (M + 1) / 2 ? ix : ix - M);
                const double t1 = mx * ux + my * uy + mz * uz;
                const double t2 = mx * vx + my * vy + mz * vz;
                if (j == 0 || ix == (M + 1) / 2) {
                    sincospi(t1, &a1i, &a1r);
                    sincospi(t1 * inv_c, &b1i, &b1r);
                    sincospi(t2, &a2i, &a2r);
                    sincospi(t2 * inv_r, &b2i, &b2r);
                } else {
                    double r, im;
                    r = a1r * su.x - a1i * su.y; im = a1r * su.y + a1i * su.x; a1r = r; a1i = im;
                    r = b1r * su.z - b1i * su.w; im = b1r * su.w + b1i * su.z; b1r = r; b1i = im;
                    r = a2r * sv.x - a2i * sv.y; im = a2r * sv.y + a2i * sv.x; a2r = r; a2i = im;
                    r = b2r * sv.z - b2i * sv.w; im = b2r * sv.w + b2i * sv.z; b2r = r; b2i = im;
                }
                // D_ncols(t1) D_nrows(t2) = (a1i a2i) / (b1i b2i) away from
                // the removable zeros.  The quotient takes the hardware
                // reciprocal refined by one Newton step (|b1i b2i| > 1e-6; the
                // step squares the estimate's relative error) instead of the
                // full division sequence.
                const bool f1 = fabs(b1i) > 1e-3, f2 = fabs(b2i) > 1e-3;
                if (f1 && f2) {
                    const double den = b1i * b2i;
                    double r = __builtin_amdgcn_rcp(den);
                    r = fma(r, fma(-den, r, 1.0), r);
                    acc[j] += (a1i * a2i) * r;
                } else {
                    const double d1 = f1 ? a1i / b1i : dirichlet_d(t1, ncols, inv_c);
                    const double d2 = f2 ? a2i / b2i : dirichlet_d(t2, nrows, inv_r);
                    acc[j] += d1 * d2;
                }
            }
        }
    }
    if (!valid) return;
    const bool mirror = !full && iz >= 1 && 2 * iz < M1;
#pragma unroll
    for (int j = 0; j < P3_RUN; ++j) {
        const int ix = ix0 + j;
        if (ix >= M) break;
        P[((size_t)iz * M + iy) * M + ix] = make_double2(acc[j], 0.0);
        if (mirror)
            P[((size_t)(M1 - iz) * M + (M - iy) % M) * M + (M - ix) % M] = make_double2(acc[j], 0.0);
    }
}

// in-place FP64 FFT of contiguous lines of length L (nl lines per block)
template <int SIGN, int NT, int MAXE>
__global__ void __launch_bounds__(NT)
k_psf3_rows(double2 *data, int64_t nlines, int nl, FftPlan plan, const double2 *__restrict__ tw) {
    double2 *buf = reinterpret_cast<double2 *>(lc_smem);
    const int L = plan.L;
    const int64_t row0 = (int64_t)blockIdx.x * nl;
    const int nloc = (int)min((int64_t)nl, nlines - row0);
    for (int idx = threadIdx.x; idx < nl * L; idx += NT) {
        const int l = idx / L, e = idx - l * L;
        buf[lds_row(l, e, L)] = (l < nloc) ? data[(row0 + l) * L + e] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, SIGN, NT, MAXE, false>(buf, nl, L, plan, tw);
    for (int idx = threadIdx.x; idx < nloc * L; idx += NT) {
        const int l = idx / L, e = idx - l * L;
        data[(row0 + l) * L + e] = buf[lds_row(l, e, L)];
    }
}

// in-place FP64 FFT along y (stride M) for columns x < ncol of every z plane;
// B columns per block.  grid (ceil(ncol / B), M1)
template <int SIGN, int NT, int MAXE>
__global__ void __launch_bounds__(NT)
k_psf3_cols(double2 *P, int M, int ncol, int B, FftPlan plan, const double2 *__restrict__ tw) {
    double2 *buf = reinterpret_cast<double2 *>(lc_smem);
    const int x0 = blockIdx.x * B;
    double2 *pl = P + (size_t)blockIdx.y * M * M;
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        buf[lds_col(c, i, B)] = (x0 + c < ncol) ? pl[(size_t)i * M + x0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, SIGN, NT, MAXE, true>(buf, B, 0, plan, tw);
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        if (x0 + c < ncol) pl[(size_t)i * M + x0 + c] = buf[lds_col(c, i, B)];
    }
}

// FP64 FFT along z of columns (k2, k3 < H), real part x scale -> khat3 in
// column blocks of width KB: value (z, k2, k3) at ((k3 / KB) M1 + z) M KB +
// k2 KB + k3 % KB (OT: float or double).  grid (ceil(H / B), M)
template <typename OT, int NT, int MAXE>
__global__ void __launch_bounds__(NT)
k_psf3_slices(const double2 *P, OT *khat3, int M1, int M, int H, int B, int KB, FftPlan plan,
              const double2 *__restrict__ tw, double scale) {
    double2 *buf = reinterpret_cast<double2 *>(lc_smem);
    const int k30 = blockIdx.x * B;
    const int k2 = blockIdx.y;
    for (int idx = threadIdx.x; idx < M1 * B; idx += NT) {
        const int z = idx / B, c = idx - z * B;
        buf[lds_col(c, z, B)] = (k30 + c < H) ? P[((size_t)z * M + k2) * M + k30 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, -1, NT, MAXE, true>(buf, B, 0, plan, tw);
    for (int idx = threadIdx.x; idx < M1 * B; idx += NT) {
        const int z = idx / B, c = idx - z * B;
        const int k3 = k30 + c;
        if (k3 < H)
            khat3[(((size_t)(k3 / KB) * M1 + z) * M + k2) * KB + k3 % KB] =
                (OT)(buf[lds_col(c, z, B)].x * scale);
    }
}

void geom3d_build_psf(Ctx &c, Geom3D &g, bool f64) {
    if (g.n2 != g.n3)
        throw Error(LC_EUNSUP, "3-D Toeplitz operator: recon dims n2 != n3 are not supported");
    const int M1 = (int)g.M1, M = (int)g.M, H = (int)g.H;
    FftPlan plan, plan1;
    LC_REQUIRE(fft_make_plan(M, plan) && fft_make_plan(M1, plan1), "3-D PSF size is not 2^p 3^q");
    const size_t np = (size_t)M1 * M * M;
    double2 *P = reinterpret_cast<double2 *>(c.scratch(0, sizeof(double2) * np));
    const int runs = (M + P3_RUN - 1) / P3_RUN;
    const int64_t threads = (int64_t)(M1 / 2 + 1) * M * runs;
    hipLaunchKernelGGL(k_psf3d_gen, dim3(ceil_div(threads, 256)), dim3(256), 0, c.stream(), P,
                       g.d_rot.as<double>(), (int)g.nproj, (int)g.nrows, (int)g.ncols, M1, M, 0);
    LC_HIP(hipGetLastError());
    constexpr int NT = 256, MAXE = 16;
    const double2 *tw = c.twiddles<double>(M);
    const double2 *tw1 = c.twiddles<double>(M1);
    // x: every row of length M
    {
        const int nl = std::max(1, NT * MAXE / M);
        const int64_t nlines = (int64_t)M1 * M;
        hipLaunchKernelGGL((k_psf3_rows<-1, NT, MAXE>), dim3(ceil_div(nlines, nl)), dim3(NT),
                           sizeof(double2) * lds_round(nl * M), c.stream(), P, nlines, nl, plan, tw);
        LC_HIP(hipGetLastError());
    }
    // y: only the columns kx < H are needed further on
    {
        const int B = std::max(1, NT * MAXE / M);
        hipLaunchKernelGGL((k_psf3_cols<-1, NT, MAXE>), dim3(ceil_div(H, B), M1), dim3(NT),
                           sizeof(double2) * lds_round(B * M), c.stream(), P, M, H, B, plan, tw);
        LC_HIP(hipGetLastError());
    }
    // z + extraction
    {
        const int B = std::max(1, std::min(NT * MAXE / M1, 64));
        const double scale = 1.0 / ((double)M1 * M * M * (double)g.N * (double)g.N);
        const size_t sh = sizeof(double2) * lds_round(B * M1);
        const dim3 grid(ceil_div(H, B), M);
        // the 3-D Toeplitz class's column block width; k3 in [H, nblk KB) stays zero
        const int KB = toep3d_block_width((int)g.N, M, f64);
        const size_t nk = (size_t)ceil_div(H, KB) * KB * M * M1;
        DBuf &kh = f64 ? g.khat3_64 : g.khat3;
        kh.alloc((f64 ? sizeof(double) : sizeof(float)) * nk);
        LC_HIP(hipMemsetAsync(kh.as(), 0, (f64 ? sizeof(double) : sizeof(float)) * nk, c.stream()));
        if (f64) {
            g.kB64 = KB;
            hipLaunchKernelGGL((k_psf3_slices<double, NT, MAXE>), grid, dim3(NT), sh, c.stream(), P,
                               g.khat3_64.as<double>(), M1, M, H, B, KB, plan1, tw1, scale);
        } else {
            g.kB = KB;
            hipLaunchKernelGGL((k_psf3_slices<float, NT, MAXE>), grid, dim3(NT), sh, c.stream(), P,
                               g.khat3.as<float>(), M1, M, H, B, KB, plan1, tw1, scale);
        }
        LC_HIP(hipGetLastError());
    }
    c.sync();
    if (f64) g.psf64_ready = true;
    else g.psf_ready = true;
}

// reference-layout kernel_hat: FP64 FFT along z of the (x, y)-transformed full
// PSF (centred layout) for k3 < H, times the phase of the reference's storage
// with the centre at h = floor(dims/2) on every axis:
// exp(-2 pi i (k1 h1 / M1 + (k2 + k3) h / M)).  grid (ceil(H / B), M)
template <typename OT, int NT, int MAXE>
__global__ void __launch_bounds__(NT)
k_khat3_export(const double2 *P, cpx<OT> *out, int M1, int M, int H, int B, FftPlan plan,
               const double2 *__restrict__ tw) {
    double2 *buf = reinterpret_cast<double2 *>(lc_smem);
    const int k30 = blockIdx.x * B;
    const int k2 = blockIdx.y;
    for (int idx = threadIdx.x; idx < M1 * B; idx += NT) {
        const int z = idx / B, c = idx - z * B;
        buf[lds_col(c, z, B)] = (k30 + c < H) ? P[((size_t)z * M + k2) * M + k30 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, -1, NT, MAXE, true>(buf, B, 0, plan, tw);
    for (int idx = threadIdx.x; idx < M1 * B; idx += NT) {
        const int c = idx / M1, k1 = idx - c * M1;
        const int k3 = k30 + c;
        if (k3 >= H) continue;
        const double2 v = buf[lds_col(c, k1, B)];
        const long long p2 = ((long long)(k2 + k3) * (M / 2)) % M;
        const long long p1 = ((long long)k1 * (M1 / 2)) % M1;
        double s, co;
        sincospi(-2.0 * ((double)p2 / (double)M + (double)p1 / (double)M1), &s, &co);
        out[((size_t)k1 * M + k2) * H + k3] =
            cmake<cpx<OT>>((OT)(v.x * co - v.y * s), (OT)(v.x * s + v.y * co));
    }
}

template <typename OT> void geom3d_kernel_hat(Ctx &c, Geom3D &g, OT *out_host) {
    if (g.n2 != g.n3)
        throw Error(LC_EUNSUP, "3-D Toeplitz operator: recon dims n2 != n3 are not supported");
    // the reference's layout and size, independent of the convolution size
    const int M1 = (int)g.M1, M = (int)g.Mr, H = M / 2 + 1;
    FftPlan plan, plan1;
    LC_REQUIRE(fft_make_plan(M, plan) && fft_make_plan(M1, plan1), "3-D PSF size is not 2^p 3^q");
    double2 *P = reinterpret_cast<double2 *>(c.scratch(0, sizeof(double2) * (size_t)M1 * M * M));
    const int runs = (M + P3_RUN - 1) / P3_RUN;
    const int64_t threads = (int64_t)M1 * M * runs;
    hipLaunchKernelGGL(k_psf3d_gen, dim3(ceil_div(threads, 256)), dim3(256), 0, c.stream(), P,
                       g.d_rot.as<double>(), (int)g.nproj, (int)g.nrows, (int)g.ncols, M1, M, 1);
    LC_HIP(hipGetLastError());
    constexpr int NT = 256, MAXE = 16;
    const double2 *tw = c.twiddles<double>(M);
    const double2 *tw1 = c.twiddles<double>(M1);
    {
        const int nl = std::max(1, NT * MAXE / M);
        const int64_t nlines = (int64_t)M1 * M;
        hipLaunchKernelGGL((k_psf3_rows<-1, NT, MAXE>), dim3(ceil_div(nlines, nl)), dim3(NT),
                           sizeof(double2) * lds_round(nl * M), c.stream(), P, nlines, nl, plan, tw);
        LC_HIP(hipGetLastError());
    }
    {
        const int B = std::max(1, NT * MAXE / M);
        hipLaunchKernelGGL((k_psf3_cols<-1, NT, MAXE>), dim3(ceil_div(H, B), M1), dim3(NT),
                           sizeof(double2) * lds_round(B * M), c.stream(), P, M, H, B, plan, tw);
        LC_HIP(hipGetLastError());
    }
    const size_t n = (size_t)M1 * M * H;
    cpx<OT> *tmp = reinterpret_cast<cpx<OT> *>(c.scratch(1, sizeof(cpx<OT>) * n));
    {
        const int B = std::max(1, std::min(NT * MAXE / M1, 64));
        hipLaunchKernelGGL((k_khat3_export<OT, NT, MAXE>), dim3(ceil_div(H, B), M), dim3(NT),
                           sizeof(double2) * lds_round(B * M1), c.stream(), P, tmp, M1, M, H, B,
                           plan1, tw1);
        LC_HIP(hipGetLastError());
    }
    LC_HIP(hipMemcpyAsync(out_host, tmp, sizeof(cpx<OT>) * n, hipMemcpyDeviceToHost, c.stream()));
    c.sync();
}
template void geom3d_kernel_hat<float>(Ctx &, Geom3D &, float *);
template void geom3d_kernel_hat<double>(Ctx &, Geom3D &, double *);

// ---------------------------------------------------------------------------
// 3-D Toeplitz middle pass
// ---------------------------------------------------------------------------
// FP32: column FFTs B * M <= 8192, slice-axis FFTs kb2 * B * M1 <= 4096.
// FP64 (16-byte complex) keeps every block within 64 KiB of LDS: 4096 / 4096.
// Column blocks: FP32 512 threads x 16 elements (M = 1536, B = 4: 12 per
// thread, 6 waves per SIMD within the 48 KiB of LDS per block), FP64 256 x 16.
// Slice-axis blocks (4096 elements): FP32 512 threads x 8, FP64 256 x 16.
template <typename R> constexpr int t3_nt() { return sizeof(R) == 4 ? 512 : 256; }
template <typename R> constexpr int t3_me() { return 16; }
template <typename R> constexpr int t3b_nt() { return sizeof(R) == 4 ? 512 : 256; }
template <typename R> constexpr int t3b_me() { return sizeof(R) == 4 ? 8 : 16; }

__device__ __forceinline__ bool prob_active3(const CgState *st, int prob) {
    return st == nullptr || st[prob].active != 0;
}

// P2f: U block (N x B, slice i1) -> FFT_M down columns -> W[kb][i1][k2][c].
// LF > 0: M = LF at compile time (unrolled stages, B = 1 << g.lsh lines), with
// the zero rows N.. of the input pruned from the first stage.
template <typename R, int NT, int MAXE, int LF>
__global__ void __launch_bounds__(NT)
k_t3_cols_fwd(const SliceRef *__restrict__ sl, const CgState *__restrict__ st,
              const cpx<R> *__restrict__ U, cpx<R> *__restrict__ W, ToepGeom g, int n1, FftPlan plan,
              const cpx<R> *__restrict__ tw) {
    using C = cpx<R>;
    const int lin = blockIdx.x + blockIdx.y * gridDim.x;
    const int logical = xcd_remap(lin, gridDim.x * gridDim.y);
    const int kb = logical / n1, i1 = logical - kb * n1;
    const SliceRef s = sl[i1];
    if (!prob_active3(st, s.prob)) return;
    const int N = g.N, M = g.M, B = g.B;
    C *buf = reinterpret_cast<C *>(lc_smem);
    const C *blk = U + s.uoff + (int64_t)kb * N * B;
    // a compiled length's pruned first stage never reads rows >= (M + 1) / 2
    const int rows = LF > 0 ? (LF + 1) / 2 : M;
    for (int idx = threadIdx.x; idx < rows * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        buf[lds_col(c, i, B)] = i < N ? blk[idx] : cmake<C>(0, 0);
    }
    __syncthreads();
    if constexpr (LF > 0) fft_lds_ct<R, LF, -1, NT, MAXE, true, -1, true, false>(buf, B, g.lsh, tw);
    else fft_lds<R, -1, NT, MAXE, true>(buf, B, 0, plan, tw);
    C *w = W + ((int64_t)kb * n1 + i1) * M * B;
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int k2 = idx / B, c = idx - k2 * B;
        w[idx] = buf[lds_col(c, k2, B)];
    }
}

// P2b: for kb2 consecutive k2 of column block kb: FFT_M1 along the slices
// (n1 values, zero padded), x khat3 (k2, k3, k1), inverse, keep n1.
// Lines (j, c) -> line j * B + c.  grid (nblk, ceil(M / kb2))
template <typename R, int NT, int MAXE>
__global__ void __launch_bounds__(NT)
k_t3_slices(const SliceRef *__restrict__ sl, const CgState *__restrict__ st, cpx<R> *__restrict__ W,
            const R *__restrict__ khat3, ToepGeom g, int n1, int M1, int kb2, FftPlan plan1,
            const cpx<R> *__restrict__ tw1) {
    using C = cpx<R>;
    if (!prob_active3(st, sl[0].prob)) return;
    const int M = g.M, B = g.B;
    const int kb = blockIdx.x, k20 = blockIdx.y * kb2;
    const int nk2 = min(kb2, M - k20);
    const int nl = kb2 * B;
    C *buf = reinterpret_cast<C *>(lc_smem);
    // slice e's values of these lines: w[e M B + line] (j * B + c = line)
    C *w = W + ((int64_t)kb * n1 * M + k20) * B;
    for (int idx = threadIdx.x; idx < nl * M1; idx += NT) {
        const int e = idx / nl, line = idx - e * nl;     // e: slice index, line = j*B + c
        const int j = line / B;
        C v = cmake<C>(0, 0);
        if (j < nk2 && e < n1) v = w[(int64_t)e * M * B + line];
        buf[lds_col(line, e, nl)] = v;
    }
    __syncthreads();
    fft_lds<R, -1, NT, MAXE, true>(buf, nl, 0, plan1, tw1);
    // the spectrum's column block kb holds these lines contiguously per slice
    // index e: kh[e M B + line]
    const R *kh = khat3 + ((int64_t)kb * M1 * M + k20) * B;
    for (int idx = threadIdx.x; idx < nl * M1; idx += NT) {
        const int e = idx / nl, line = idx - e * nl;
        const int j = line / B;
        const R kv = j < nk2 ? kh[(int64_t)e * M * B + line] : (R)0;
        C *p = buf + lds_col(line, e, nl);
        *p = cscale(*p, kv);
    }
    __syncthreads();
    fft_lds<R, 1, NT, MAXE, true>(buf, nl, 0, plan1, tw1);
    for (int idx = threadIdx.x; idx < nl * n1; idx += NT) {
        const int e = idx / nl, line = idx - e * nl;
        const int j = line / B;
        if (j < nk2) w[(int64_t)e * M * B + line] = buf[lds_col(line, e, nl)];
    }
}

// P2b with a compiled slice-axis length L1 (a power of two): NL = NT MAXE / L1
// lines per block (kb2 = NL / B rows k2), only the n1 <= (L1 + 1) / 2 leading
// rows loaded, and the forward FFT, the khat3 multiply and the inverse FFT run
// as one fused convolution (fft_conv_ct).  grid (nblk, ceil(M / kb2))
template <typename R, int NT, int MAXE, int L1>
__global__ void __launch_bounds__(NT)
k_t3_slices_ct(const SliceRef *__restrict__ sl, const CgState *__restrict__ st, cpx<R> *__restrict__ W,
               const R *__restrict__ khat3, ToepGeom g, int n1, int kb2,
               const cpx<R> *__restrict__ tw1) {
    using C = cpx<R>;
    using CM = ConvMid<R, L1, MAXE>;
    constexpr int NL = NT * MAXE / L1;
    constexpr int LSH = NL == 1 ? 0 : NL == 2 ? 1 : NL == 4 ? 2 : NL == 8 ? 3 : NL == 16 ? 4 : NL == 32 ? 5
                      : NL == 64 ? 6 : NL == 128 ? 7 : 8;
    static_assert((1 << LSH) == NL, "lines per block must be a power of two up to 256");
    constexpr int HALF = (L1 + 1) / 2;
    if (!prob_active3(st, sl[0].prob)) return;
    const int M = g.M, B = g.B, bsh = g.lsh;
    const int kb = blockIdx.x, k20 = blockIdx.y * kb2;
    const int nk2 = min(kb2, M - k20);
    C *buf = reinterpret_cast<C *>(lc_smem);
    // slice e's values of these lines: w[e M B + line] (lines j << bsh | c
    // are consecutive values); the spectrum's column block kb holds them at
    // kh[e M B + line] (neighbouring lines of a wave read neighbouring values)
    C *w = W + ((int64_t)kb * n1 * M + k20) * B;
    const R *kh = khat3 + ((int64_t)kb * L1 * M + k20) * B;
    // spectrum values of this thread's middle-stage butterflies (line, index)
    R kr[CM::MAXB][CM::R];
#pragma unroll
    for (int t = 0; t < CM::MAXB; ++t) {
        int line = 0, jm = 0;
        const bool has = conv_mid_index<R, L1, NT, MAXE, LSH>(t, NL, LSH, line, jm);
        const bool ok = has && (line >> bsh) < nk2;
#pragma unroll
        for (int r = 0; r < CM::R; ++r)
            kr[t][r] = ok ? kh[(int64_t)(jm + r * CM::LR) * M * B + line] : (R)0;
    }
    for (int idx = threadIdx.x; idx < NL * HALF; idx += NT) {
        const int e = idx >> LSH, line = idx & (NL - 1);
        const int j = line >> bsh;
        C v = cmake<C>(0, 0);
        if (j < nk2 && e < n1) v = w[(int64_t)e * M * B + line];
        buf[swz((e << LSH) + line)] = v;
    }
    __syncthreads();
    fft_conv_ct<R, L1, NT, MAXE, LSH>(buf, NL, LSH, tw1, kr);
    for (int idx = threadIdx.x; idx < NL * n1; idx += NT) {
        const int e = idx >> LSH, line = idx & (NL - 1);
        const int j = line >> bsh;
        if (j < nk2) w[(int64_t)e * M * B + line] = buf[swz((e << LSH) + line)];
    }
}

// P2i: W[kb][i1][:][:] -> inverse FFT_M down columns -> keep N rows -> U block
// (LF > 0: compiled length; the unused outputs N.. pruned from the last stage)
template <typename R, int NT, int MAXE, int LF>
__global__ void __launch_bounds__(NT)
k_t3_cols_inv(const SliceRef *__restrict__ sl, const CgState *__restrict__ st,
              const cpx<R> *__restrict__ W, cpx<R> *__restrict__ U, ToepGeom g, int n1, FftPlan plan,
              const cpx<R> *__restrict__ tw) {
    using C = cpx<R>;
    const int lin = blockIdx.x + blockIdx.y * gridDim.x;
    const int logical = xcd_remap(lin, gridDim.x * gridDim.y);
    const int kb = logical / n1, i1 = logical - kb * n1;
    const SliceRef s = sl[i1];
    if (!prob_active3(st, s.prob)) return;
    const int N = g.N, M = g.M, B = g.B;
    C *buf = reinterpret_cast<C *>(lc_smem);
    const C *w = W + ((int64_t)kb * n1 + i1) * M * B;
    for (int idx = threadIdx.x; idx < M * B; idx += NT) {
        const int k2 = idx / B, c = idx - k2 * B;
        buf[lds_col(c, k2, B)] = w[idx];
    }
    __syncthreads();
    if constexpr (LF > 0) fft_lds_ct<R, LF, 1, NT, MAXE, true, -1, false, true>(buf, B, g.lsh, tw);
    else fft_lds<R, 1, NT, MAXE, true>(buf, B, 0, plan, tw);
    C *blk = U + s.uoff + (int64_t)kb * N * B;
    for (int idx = threadIdx.x; idx < N * B; idx += NT) {
        const int i = idx / B, c = idx - i * B;
        blk[idx] = buf[lds_col(c, i, B)];
    }
}

template <typename R> static const cpx<R> *t3_tw1(const ToepClass &k);
template <> const float2 *t3_tw1<float>(const ToepClass &k) { return k.tw1; }
template <> const double2 *t3_tw1<double>(const ToepClass &k) { return k.tw1_64; }
template <typename R> static const R *t3_khat(const Geom3D &g);
template <> const float *t3_khat<float>(const Geom3D &g) { return g.khat3.as<float>(); }
template <> const double *t3_khat<double>(const Geom3D &g) { return g.khat3_64.as<double>(); }

template <typename R> void toep3d_mid(Ctx &c, const ToepClass &k, const CgState *st, cpx<R> *U) {
    using C = cpx<R>;
    constexpr int NT = t3_nt<R>(), ME = t3_me<R>(), NTB = t3b_nt<R>(), MEB = t3b_me<R>();
    const bool f64 = sizeof(R) == 8;
    LC_REQUIRE(k.B * k.M <= toep3d_col_cap(f64) && k.kb2 * k.B * k.M1 <= toep3d_slice_cap(f64),
               "toeplitz 3-D: block exceeds the LDS shape");
    const SliceRef *sl = k.slices.as<SliceRef>();
    C *W = k.W.as<C>();
    const C *tw = toep_tw<R>(k);
    const size_t shc = sizeof(C) * lds_round(k.B * k.M);
    // FP32 column FFTs at the Toeplitz kernels' compiled lengths, runtime plans otherwise
    const int lf = f64 ? 0 : k.M;
    auto cols = [&](auto dir) {
        constexpr bool FWD = decltype(dir)::value;
#define LC_T3_COLS(L)                                                                             \
    if constexpr (FWD)                                                                            \
        hipLaunchKernelGGL((k_t3_cols_fwd<R, NT, ME, L>), dim3(k.nblk, k.n1), dim3(NT), shc,       \
                           c.stream(), sl, st, U, W, k.g, k.n1, k.plan, tw);                      \
    else                                                                                          \
        hipLaunchKernelGGL((k_t3_cols_inv<R, NT, ME, L>), dim3(k.nblk, k.n1), dim3(NT), shc,       \
                           c.stream(), sl, st, W, U, k.g, k.n1, k.plan, tw)
        switch (lf) {
        case 256: LC_T3_COLS(256); break;
        case 384: LC_T3_COLS(384); break;
        case 512: LC_T3_COLS(512); break;
        case 768: LC_T3_COLS(768); break;
        case 1024: LC_T3_COLS(1024); break;
        case 1536: LC_T3_COLS(1536); break;
        case 2048: LC_T3_COLS(2048); break;
        default: LC_T3_COLS(0); break;
        }
#undef LC_T3_COLS
        LC_HIP(hipGetLastError());
    };
    cols(std::true_type{});
    const size_t shb = sizeof(C) * lds_round(k.kb2 * k.B * k.M1);
    const dim3 gb(k.nblk, ceil_div(k.M, k.kb2));
    // FP32 slice-axis lengths 16..128 with a full block of lines: compiled,
    // fused convolution; anything else: runtime plan
    const bool ct1 = !f64 && k.kb2 * k.B * k.M1 == NTB * MEB &&
                     (k.M1 == 16 || k.M1 == 32 || k.M1 == 64 || k.M1 == 128);
#define LC_T3_SL(L1)                                                                              \
    hipLaunchKernelGGL((k_t3_slices_ct<R, NTB, MEB, L1>), gb, dim3(NTB), shb, c.stream(),          \
                       sl, st, W, t3_khat<R>(*k.g3), k.g, k.n1, k.kb2, t3_tw1<R>(k))
    if (ct1 && k.M1 == 16) LC_T3_SL(16);
    else if (ct1 && k.M1 == 32) LC_T3_SL(32);
    else if (ct1 && k.M1 == 64) LC_T3_SL(64);
    else if (ct1 && k.M1 == 128) LC_T3_SL(128);
    else
        hipLaunchKernelGGL((k_t3_slices<R, NTB, MEB>), gb, dim3(NTB), shb, c.stream(), sl, st, W,
                           t3_khat<R>(*k.g3), k.g, k.n1, k.M1, k.kb2, k.plan1, t3_tw1<R>(k));
#undef LC_T3_SL
    LC_HIP(hipGetLastError());
    cols(std::false_type{});
}
template void toep3d_mid<float>(Ctx &, const ToepClass &, const CgState *, float2 *);
template void toep3d_mid<double>(Ctx &, const ToepClass &, const CgState *, double2 *);

// shape limits of the 3-D middle pass for the class set-up
int toep3d_col_cap(bool f64) { return f64 ? t3_nt<double>() * t3_me<double>() : t3_nt<float>() * t3_me<float>(); }
int toep3d_slice_cap(bool f64) {
    // 4096 complex: 32 KiB (FP32) or 64 KiB (FP64) of LDS
    return f64 ? t3b_nt<double>() * t3b_me<double>() : t3b_nt<float>() * t3b_me<float>();
}

} // namespace lc

// C++ API over the C ABI (include/laminocam_tomocam.hpp), exercised the way
// the reference's own tests exercise its API:
//   tests/test_adjoint2d.cpp        <forward2d y, z> == <y... , backproj2d z>
//   tests/test_sysmat_toeplitz2d.cpp sysmat2d(x, grid) == sysmat2d(x, psf)
//   tests/test_sysmat_toeplitz.cpp   sysmat(x, grid) == sysmat(x, psf)
//   tests/test_cg_normal.cpp         CG on A x = A x_true reduces the residual
//   src/mbir2d.cpp:66-100, src/mbir.cpp:134  solvers over opaque lambdas (host loop)
//   src/projection.cpp:105-108       shift count mismatch -> std::invalid_argument
//   include/mbir2d_prep.h:44-65      prepare_sinogram: normalised, {rows, angles, cols}
//   src/mbir.cpp:34-161              MBIR == the same steps through the API, cropped
// Seeded inputs from a SplitMix64 stream; exits nonzero on the first failure.
#define LAMINOCAM_STANDALONE_ARRAY
#include "laminocam_tomocam.hpp"

#include <cmath>
#include <cstdio>

using namespace tomocam;

static uint64_t g_state = 0x1A41C0DEull;
static double urand() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T> static Array<T> random_array(dims_t d, double lo = 0.0, double hi = 1.0) {
    Array<T> a(d);
    for (size_t i = 0; i < a.size(); ++i) a[i] = (T)(lo + (hi - lo) * urand());
    return a;
}

template <typename T> static double dot(const Array<T> &a, const Array<T> &b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += (double)a[i] * (double)b[i];
    return s;
}

template <typename T> static double rel_l2(const Array<T> &a, const Array<T> &b) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        const double d = (double)a[i] - (double)b[i];
        num += d * d;
        den += (double)b[i] * (double)b[i];
    }
    return std::sqrt(num / den);
}

static int g_fail = 0;
#define CHECK(name, cond)                                                     \
    do {                                                                      \
        const bool ok_ = (cond);                                              \
        std::printf("%s %s\n", ok_ ? "PASS" : "FAIL", name);                  \
        std::fflush(stdout);                                                  \
        if (!ok_) ++g_fail;                                                   \
    } while (0)

int main() {
    std::printf("[case] C++ API over the C ABI\n");
    std::fflush(stdout);
    // ---- 2-D ----
    {
        const size_t N = 64, nang = 90;
        std::vector<float> th(nang);
        for (size_t i = 0; i < nang; ++i) th[i] = (float)(((double)i - nang / 2.0) * M_PI / nang);
        cpu::PolarGrid2D<float> grid(th, N);
        Array<float> x = random_array<float>(dims_t{2, N, N});
        Array<float> z = random_array<float>(dims_t{2, nang, N}, -1.0, 1.0);
        Array<float> fx = forward2d(x, grid);
        Array<float> bz = backproj2d(z, grid, dims_t{2, N, N});
        const double l = dot(fx, z), r = dot(x, bz);
        CHECK("adjoint2d", std::fabs(l - r) <= 1e-5 * (std::fabs(l) + std::fabs(r)));
        cpu::PointSpreadFunction2D<float> psf(grid, N);
        CHECK("psf_size", psf.psf_size() == 128);
        Array<float> t = sysmat2d(x, psf);
        Array<float> s = sysmat2d(x, grid);
        CHECK("sysmat2d toeplitz == nufft", rel_l2(t, s) < 1e-4);
        Array<std::complex<float>> kh = psf.kernel_hat();
        CHECK("kernel_hat dims", kh.dims() == (dims_t{1, 128, 65}));
        // CG on the normal equations (tests/test_cg_normal.cpp)
        Array<float> xt = random_array<float>(dims_t{1, N, N});
        Array<float> y = sysmat2d(xt, psf);
        opt::Function<float> A = opt::toeplitz(psf);
        Array<float> x0(dims_t{1, N, N});
        Array<float> x5 = opt::cgsolver(A, y, x0, 5, 1e-12f, 0.0f);
        Array<float> x40 = opt::cgsolver(A, y, x0, 40, 1e-12f, 0.0f);
        auto resid = [&](const Array<float> &xx) {
            Array<float> ax = sysmat2d(xx, psf);
            double n = 0.0, d = 0.0;
            for (size_t i = 0; i < ax.size(); ++i) {
                n += ((double)ax[i] - y[i]) * ((double)ax[i] - y[i]);
                d += (double)y[i] * y[i];
            }
            return std::sqrt(n / d);
        };
        const double r5 = resid(x5), r40 = resid(x40);
        std::printf("  cg residual 5 it %.3e, 40 it %.3e\n", r5, r40);
        CHECK("cgsolver residual decreases", r40 < r5 && r40 < 1e-2);
        // the reference's callers build the operator as an opaque lambda
        // (src/mbir2d.cpp:73-75, src/mbir.cpp:134): solved by the host loop
        opt::Function<float> opaque = [&psf](const Array<float> &v) { return sysmat2d(v, psf); };
        Array<float> x5o = opt::cgsolver(opaque, y, x0, 5, 1e-12f, 0.0f);
        std::printf("  cg 5 it opaque vs toeplitz %.3e\n", rel_l2(x5o, x5));
        CHECK("cgsolver opaque callable == toeplitz", rel_l2(x5o, x5) < 1e-5);
        Array<float> sb = opt::split_bregman(A, y, x0, 0.1f, 10.0f, 3, 1, 1e-12f, 0.0f);
        CHECK("split_bregman runs", std::isfinite(dot(sb, sb)));
        // Split-Bregman on a 5-slice stack, where the interior differences
        // and so the shrink are active: host loop around the opaque callable
        // against the device solve
        Array<float> xs5 = random_array<float>(dims_t{5, N, N});
        Array<float> y5 = sysmat2d(xs5, psf);
        Array<float> z5(dims_t{5, N, N});
        Array<float> sbd = opt::split_bregman(A, y5, z5, 0.05f, 10.0f, 4, 2, 1e-12f, 0.0f);
        Array<float> sbo = opt::split_bregman(opaque, y5, z5, 0.05f, 10.0f, 4, 2, 1e-12f, 0.0f);
        std::printf("  split_bregman 4x2 opaque vs toeplitz %.3e\n", rel_l2(sbo, sbd));
        CHECK("split_bregman opaque callable == toeplitz", rel_l2(sbo, sbd) < 1e-4);
        // the flow of cpu_mbir2d (src/mbir2d.cpp:66-100) written against this
        // API: grid, PSF, opaque operator, back-projected rhs, solve from zeros;
        // against the device driver gpu::MBIR2D on the same sinogram
        {
            Array<float> sino = forward2d(xs5, grid);
            ReconParams prm;
            prm.recon_dims = {5, N, N};
            prm.maxIters = 6;
            prm.innerIters = 2;
            for (int reg = 0; reg < 2; ++reg) {
                prm.regularizer = reg ? Regularizer::SPLIT_BREGMAN : Regularizer::UNCONSTRAINED;
                const dims_t recon_dims{prm.recon_dims[0], prm.recon_dims[1], prm.recon_dims[2]};
                cpu::PolarGrid2D<float> pg(th, sino.ncols());
                cpu::PointSpreadFunction2D<float> P(pg, recon_dims.n3);
                opt::Function<float> Aop = [&P](const Array<float> &v) { return sysmat2d(v, P); };
                Array<float> rhs = backproj2d(sino, pg, recon_dims);
                Array<float> xz = Array<float>::zeros(recon_dims);
                Array<float> rec = reg ? opt::split_bregman<float>(Aop, rhs, xz, prm.lambda, prm.mu, prm.maxIters,
                                                                    prm.innerIters, prm.tol, prm.xtol)
                                       : opt::cgsolver<float>(Aop, rhs, xz, prm.maxIters, prm.tol, prm.xtol);
                Array<float> drv = gpu::MBIR2D(sino, th, prm);
                std::printf("  cpu_mbir2d flow (%s) vs MBIR2D %.3e\n", reg ? "SB" : "CG", rel_l2(rec, drv));
                CHECK(reg ? "cpu_mbir2d flow SB == MBIR2D" : "cpu_mbir2d flow CG == MBIR2D", rel_l2(rec, drv) < 1e-4);
            }
        }
        // interior-only differences (src/finitediff.cpp:26-90) need n1 >= 3
        Array<float> v = random_array<float>(dims_t{5, 32, 32});
        auto g = opt::grad_u(v);
        Array<float> lap = opt::laplacian(v);
        Array<float> dv = opt::divergence(g);
        CHECK("laplacian == divergence(grad)", dot(lap, lap) > 0.0 && rel_l2(lap, dv) < 1e-6);
        // double arrays are converted at the boundary
        std::vector<double> thd(th.begin(), th.end());
        cpu::PolarGrid2D<double> gridd(thd, N);
        Array<double> xd(dims_t{1, N, N});
        for (size_t i = 0; i < xd.size(); ++i) xd[i] = x[i];
        Array<double> fd = forward2d(xd, gridd);
        double md = 0.0;
        for (size_t i = 0; i < fd.size(); ++i) md = std::fmax(md, std::fabs(fd[i] - fx[i]));
        CHECK("double arrays", md < 1e-5);
        // Array<double> runs the FP64 Toeplitz operator, kernel_hat, solvers
        // and differences: A^T A is symmetric to FP64 rounding
        cpu::PointSpreadFunction2D<double> psfd(gridd, N);
        Array<double> ud = random_array<double>(dims_t{2, N, N}), wd = random_array<double>(dims_t{2, N, N});
        Array<double> aud = sysmat2d(ud, psfd), awd = sysmat2d(wd, psfd);
        const double ld = dot(aud, wd), rd_ = dot(ud, awd);
        std::printf("  fp64 toeplitz symmetry %.3e\n", std::fabs(ld - rd_) / std::fabs(ld));
        CHECK("fp64 toeplitz symmetric", std::fabs(ld - rd_) <= 1e-12 * std::fabs(ld));
        Array<std::complex<double>> khd = psfd.kernel_hat();
        double kdiff = 0.0, knorm = 0.0;
        for (size_t i = 0; i < khd.size(); ++i) {
            kdiff += std::norm(khd[i] - std::complex<double>(kh[i]));
            knorm += std::norm(khd[i]);
        }
        CHECK("fp64 kernel_hat", std::sqrt(kdiff / knorm) < 1e-6);
        Array<double> yd(dims_t{1, N, N});
        for (size_t i = 0; i < yd.size(); ++i) yd[i] = y[i];
        Array<double> x0d(dims_t{1, N, N});
        Array<double> x5d = opt::cgsolver(opt::toeplitz(psfd), yd, x0d, 5, 1e-12, 0.0);
        double cdiff = 0.0, cnorm = 0.0;
        for (size_t i = 0; i < x5d.size(); ++i) {
            cdiff += (x5d[i] - x5[i]) * (x5d[i] - x5[i]);
            cnorm += x5d[i] * x5d[i];
        }
        std::printf("  cg 5 it fp64 vs fp32 %.3e\n", std::sqrt(cdiff / cnorm));
        CHECK("fp64 cgsolver", std::sqrt(cdiff / cnorm) < 1e-4);
        Array<double> vd = random_array<double>(dims_t{5, 32, 32});
        auto gd = opt::grad_u(vd);
        CHECK("fp64 laplacian == divergence(grad)",
              rel_l2(opt::laplacian(vd), opt::divergence(gd)) < 1e-14);
    }
    // ---- 3-D ----
    {
        const size_t nproj = 9, nrows = 16, ncols = 16;
        std::vector<float> th(nproj), ga(nproj, (float)(-30.0 * M_PI / 180.0)), be(nproj, 0.0f);
        for (size_t i = 0; i < nproj; ++i) th[i] = (float)(2.0 * M_PI * i / nproj);
        cpu::PolarGrid<float> grid(th, ga, be, nrows, ncols);
        const dims_t rd{5, 16, 16};
        Array<float> x = random_array<float>(rd);
        cpu::PointSpreadFunction<float> psf(grid, rd);
        Array<float> t = sysmat(x, psf);
        Array<float> s = sysmat(x, grid);
        CHECK("sysmat3d toeplitz == nufft", rel_l2(t, s) < 1e-4);
        Array<float> p = forward(x, grid);
        CHECK("forward3d dims", p.dims() == grid.dims());
        Array<float> z = random_array<float>(grid.dims(), -1.0, 1.0);
        const double l = dot(p, z), r = dot(x, backproj(z, grid, rd));
        CHECK("adjoint3d", std::fabs(l - r) <= 1e-5 * (std::fabs(l) + std::fabs(r)));
        std::vector<std::array<float, 2>> shifts(nproj - 1);
        bool threw = false;
        try {
            backproj(z, grid, rd, shifts);
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        CHECK("backproj shift count mismatch", threw);
        Array<float> y = sysmat(x, psf);
        Array<float> x3 = opt::cgsolver(opt::toeplitz(psf), y, Array<float>(rd), 4, 1e-12f, 0.0f);
        CHECK("cgsolver 3-D runs", std::isfinite(dot(x3, x3)));
        // the operator of src/mbir.cpp:134 as an opaque lambda
        opt::Function<float> A3 = [&psf](const Array<float> &v) { return sysmat(v, psf); };
        Array<float> x3o = opt::cgsolver<float>(A3, y, Array<float>::zeros(rd), 4, 1e-12f, 0.0f);
        std::printf("  cg 3-D 4 it opaque vs toeplitz %.3e\n", rel_l2(x3o, x3));
        CHECK("cgsolver 3-D opaque callable == toeplitz", rel_l2(x3o, x3) < 1e-5);
        Array<float> s3d = opt::split_bregman<float>(opt::toeplitz(psf), y, Array<float>::zeros(rd), 0.05f, 10.0f,
                                                     3, 2, 1e-12f, 0.0f);
        Array<float> s3o = opt::split_bregman<float>(A3, y, Array<float>::zeros(rd), 0.05f, 10.0f, 3, 2, 1e-12f,
                                                     0.0f);
        std::printf("  split_bregman 3-D opaque vs toeplitz %.3e\n", rel_l2(s3o, s3d));
        CHECK("split_bregman 3-D opaque callable == toeplitz", rel_l2(s3o, s3d) < 1e-4);
    }
    // ---- drivers ----
    {
        // two datasets of 5 and 4 projections {n, 3 rows, 16 cols}
        std::vector<Dataset_t<float>> ds(2);
        const size_t nr = 3, nc = 16, na[2] = {5, 4};
        float mx = 0.0f;
        for (int d = 0; d < 2; ++d) {
            ds[d].projs = random_array<float>(dims_t{na[d], nr, nc}, 0.1, 2.0);
            for (size_t i = 0; i < ds[d].projs.size(); ++i) mx = std::max(mx, ds[d].projs[i]);
            for (size_t i = 0; i < na[d]; ++i) {
                ds[d].angles.push_back((float)(2.0 * M_PI * i / na[d] + 0.3 * d));
                ds[d].shifts.push_back({(float)(urand() - 0.5), (float)(urand() - 0.5)});
            }
            ds[d].gamma = (float)(-30.0 * M_PI / 180.0);
        }
        Sinogram2D<float> sg = prepare_sinogram(ds);
        bool ok = (sg.data.dims() == dims_t{nr, 9, nc}) && sg.theta.size() == 9;
        for (size_t r = 0; r < nr && ok; ++r)
            for (size_t a = 0; a < 9 && ok; ++a)
                for (size_t c = 0; c < nc; ++c) {
                    const Array<float> &pj = a < 5 ? ds[0].projs : ds[1].projs;
                    const size_t aa = a < 5 ? a : a - 5;
                    ok = ok && sg.data[(r * 9 + a) * nc + c] == pj[(aa * nr + r) * nc + c] / mx;
                }
        CHECK("prepare_sinogram", ok);

        ReconParams rp;
        rp.recon_dims = {3, 12, 12};
        rp.maxIters = 4;
        ReconParams rp2 = rp;
        rp2.recon_dims = {3, 16, 16};
        Array<float> rec2 = gpu::MBIR2D(sg.data, sg.theta, rp2);
        CHECK("MBIR2D runs", (rec2.dims() == dims_t{3, 16, 16}) && std::isfinite(dot(rec2, rec2)));

        // MBIR against its steps: normalise, pad 3 x 16 -> 3 x 21 (int(16 * 1.4142) = 22
        // -> 21; int(3 * 1.4142) = 4 -> 3), grid, backproj with shifts into the padded
        // volume 3 x 15 x 15, CG from zeros, symmetric crop to 3 x 12 x 12
        Array<float> rec = MBIR(ds, rp);
        const size_t n3p = 21;
        const dims_t odims{3, 15, 15};
        Array<float> padded(dims_t{9, nr, n3p});
        std::vector<float> th, ga(9, ds[0].gamma), be(9, 0.0f);
        std::vector<std::array<float, 2>> sh;
        for (int d = 0; d < 2; ++d) {
            th.insert(th.end(), ds[d].angles.begin(), ds[d].angles.end());
            sh.insert(sh.end(), ds[d].shifts.begin(), ds[d].shifts.end());
        }
        for (size_t a = 0; a < 9; ++a)
            for (size_t r = 0; r < nr; ++r)
                for (size_t c = 0; c < nc; ++c) {
                    const Array<float> &pj = a < 5 ? ds[0].projs : ds[1].projs;
                    const size_t aa = a < 5 ? a : a - 5;
                    padded[(a * nr + r) * n3p + c + 2] = pj[(aa * nr + r) * nc + c] / mx;
                }
        cpu::PolarGrid<float> pg(th, ga, be, nr, n3p);
        Array<float> rhs = backproj(padded, pg, odims, sh);
        cpu::PointSpreadFunction<float> pp(pg, odims);
        Array<float> xs = opt::cgsolver(opt::toeplitz(pp), rhs, Array<float>(odims), 4, 1e-5f, 1e-5f);
        bool same = (rec.dims() == dims_t{3, 12, 12});
        for (size_t i = 0; i < 3 && same; ++i)
            for (size_t j = 0; j < 12; ++j)
                for (size_t k = 0; k < 12; ++k)
                    same = same && rec[(i * 12 + j) * 12 + k] == xs[(i * 15 + j + 1) * 15 + k + 1];
        CHECK("MBIR == its steps, cropped", same);
        std::vector<std::array<float, 2>> short_sh(3);
        std::vector<Dataset_t<float>> bad(1);
        bad[0].projs = random_array<float>(dims_t{4, nr, nc});
        bad[0].angles.assign(4, 0.0f);
        bad[0].shifts = short_sh;
        bool threw = false;
        try {
            MBIR(bad, rp);
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        CHECK("MBIR shift count mismatch", threw);
    }
    std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}

// Parts of the C++ API (include/laminocam_tomocam.hpp) that run without a
// device, checked on the CPU:
//   include/optimize.h:31-61        opt::Function is the std::function alias;
//                                   OptimizerConfig defaults
//   tests/test_cg_normal.cpp        cgsolver over an opaque callable solves the
//                                   normal equations A^T A x = A^T c of a random
//                                   sparse overdetermined A (host loop,
//                                   src/conjgrad.cpp:33-76)
//   src/conjgrad.cpp:53-56          the |<p,Ap>| < 1e-10 guard stops a zero operator
//   include/polar_grid2d.h:50-65    PolarGrid2D::npts, x, y
//   include/config.h:284-343        broadcast_shifts, apply_alignment
//   include/polar_grid.h:52-94      PolarGrid::npts, x, y, z, w against the full
//                                   rotation R^T = (Rx(theta) Ry(beta) Rz(gamma))^T
// Seeded inputs from a SplitMix64 stream; exits nonzero on the first failure.
#define LAMINOCAM_STANDALONE_ARRAY
#include "laminocam_tomocam.hpp"

#include <cstdio>
#include <tuple>
#include <type_traits>

using namespace tomocam;

static_assert(std::is_same<opt::Function<float>, std::function<Array<float>(const Array<float> &)>>::value,
              "opt::Function must be the std::function alias");
static_assert(std::is_same<opt::Residual<double>, std::function<double(const Array<double> &)>>::value,
              "opt::Residual must be the std::function alias");

static uint64_t g_state = 0x1A41C0DEull;
static double urand() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static int g_fail = 0;
#define CHECK(name, cond)                                                     \
    do {                                                                      \
        const bool ok_ = (cond);                                              \
        std::printf("%s %s\n", ok_ ? "PASS" : "FAIL", name);                  \
        std::fflush(stdout);                                                  \
        if (!ok_) ++g_fail;                                                   \
    } while (0)

// CG on A^T A x = A^T c with A a 64 x 32 sparse matrix (about 30% filled,
// entries U[-1, 1]) and c = A x_true; the solve must recover x_true.
template <typename T> static double cg_normal(T tol) {
    const size_t m = 64, n = 32;
    std::vector<std::tuple<size_t, size_t, T>> coo;
    for (size_t r = 0; r < m; ++r)
        for (size_t c = 0; c < n; ++c)
            if (urand() < 0.30) coo.emplace_back(r, c, (T)(2.0 * urand() - 1.0));
    Array<T> x_true(dims_t{1, 1, n});
    for (size_t i = 0; i < n; ++i) x_true[i] = (T)(2.0 * urand() - 1.0);
    auto A = [&](const Array<T> &v) {
        Array<T> t = Array<T>::zeros(dims_t{1, 1, m});
        for (const auto &[r, c, a] : coo) t[r] += a * v[c];
        Array<T> z = Array<T>::zeros(dims_t{1, 1, n});
        for (const auto &[r, c, a] : coo) z[c] += a * t[r];
        return z;
    };
    Array<T> b = A(x_true);
    opt::Function<T> F = A;
    Array<T> x = opt::cgsolver(F, b, Array<T>::zeros(dims_t{1, 1, n}), 2000, tol, tol / T(100));
    return (double)(array::norm2(x - x_true) / array::norm2(x_true));
}

int main() {
    std::printf("[case] C++ API, host-only parts\n");
    std::fflush(stdout);

    const double e64 = cg_normal<double>(1e-5);
    std::printf("  cg normal equations double rel_err %.3e\n", e64);
    CHECK("cgsolver<double> opaque callable, normal equations", e64 < 1e-3);
    const double e32 = cg_normal<float>(1e-3f);
    std::printf("  cg normal equations float rel_err %.3e\n", e32);
    CHECK("cgsolver<float> opaque callable, normal equations", e32 < 1e-1);

    // diagonal SPD system with condition number 100: CG is exact in as many
    // iterations as there are distinct eigenvalues (here 8)
    {
        const size_t n = 64;
        Array<double> diag(dims_t{1, 1, n}), xt(dims_t{1, 1, n});
        for (size_t i = 0; i < n; ++i) {
            diag[i] = 1.0 + 99.0 * (double)(i % 8) / 7.0;
            xt[i] = urand() - 0.5;
        }
        opt::Function<double> D = [&diag](const Array<double> &v) { return v * diag; };
        Array<double> y = D(xt);
        Array<double> x = opt::cgsolver(D, y, Array<double>::zeros(dims_t{1, 1, n}), 8, 0.0, 0.0);
        const double e = array::norm2(x - xt) / array::norm2(xt);
        std::printf("  cg diagonal kappa=100, 8 distinct eigenvalues, 8 it: rel_err %.3e\n", e);
        CHECK("cgsolver exact after 8 iterations on 8 eigenvalues", e < 1e-9);
        opt::Function<double> Z = [](const Array<double> &v) { return v * 0.0; };
        Array<double> x0 = Array<double>::ones(dims_t{1, 1, n});
        Array<double> xz = opt::cgsolver(Z, y, x0, 5, 0.0, 0.0);
        CHECK("zero operator: the pAp guard returns x0", array::norm2(xz - x0) == 0.0);
        bool threw = false;
        try {
            opt::cgsolver(opt::Function<double>(), y, x0, 5, 0.0, 0.0);
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        CHECK("empty operator -> std::invalid_argument", threw);
    }

    {
        opt::OptimizerConfig<float> cfg;
        CHECK("OptimizerConfig defaults (include/optimize.h:52-54)",
              cfg.method == opt::Optimizer::SPLIT_BREGMAN && cfg.lambda == 0.1f && cfg.mu == 5.0f &&
                  cfg.outer_max == 50 && cfg.inner_max == 4 && cfg.tol == 1e-5f && cfg.xtol == 1e-5f);
    }

    // ---- polar grid points ----
    {
        const size_t nang = 7, nrad = 10;
        std::vector<double> th(nang);
        for (size_t i = 0; i < nang; ++i) th[i] = M_PI * (double)i / nang - 0.2;
        cpu::PolarGrid2D<double> g(th, nrad);
        double md = 0.0;
        for (size_t i = 0; i < nang; ++i)
            for (size_t k = 0; k < nrad; ++k) {
                const double r = -M_PI + 2.0 * M_PI * (k + 0.5) / nrad;
                md = std::fmax(md, std::fabs(g.x[i * nrad + k] - r * std::cos(th[i])));
                md = std::fmax(md, std::fabs(g.y[i * nrad + k] - r * std::sin(th[i])));
            }
        CHECK("PolarGrid2D points", g.npts == nang * nrad && (g.x.dims() == dims_t{1, nang, nrad}) && md < 1e-14);
    }
    {
        const size_t np = 5, nr = 6, nc = 9;
        std::vector<double> th(np), ga(np), be(np);
        for (size_t i = 0; i < np; ++i) {
            th[i] = 2.0 * M_PI * (double)i / np;
            ga[i] = -0.52 + 0.1 * (double)i;
            be[i] = 0.05 * (double)i;
        }
        cpu::PolarGrid<double> g(th, ga, be, nr, nc);
        double md = 0.0;
        size_t masked = 0, masked_ref = 0;
        for (size_t i = 0; i < np; ++i) {
            // R = Rx(theta) Ry(beta) Rz(gamma), points R^T (qX, qY, 0)
            const double ct = std::cos(th[i]), st = std::sin(th[i]);
            const double cb = std::cos(be[i]), sb = std::sin(be[i]);
            const double cg = std::cos(ga[i]), sg = std::sin(ga[i]);
            const double Rx[3][3] = {{1, 0, 0}, {0, ct, -st}, {0, st, ct}};
            const double Ry[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
            const double Rz[3][3] = {{cg, -sg, 0}, {sg, cg, 0}, {0, 0, 1}};
            double T1[3][3] = {}, R[3][3] = {};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) T1[a][b] += Ry[a][c] * Rz[c][b];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) R[a][b] += Rx[a][c] * T1[c][b];
            for (size_t j = 0; j < nr; ++j)
                for (size_t k = 0; k < nc; ++k) {
                    const double qX = -M_PI + 2.0 * M_PI * (k + 0.5) / nc, qY = -M_PI + 2.0 * M_PI * (j + 0.5) / nr;
                    double q[3];
                    for (int a = 0; a < 3; ++a) q[a] = R[0][a] * qX + R[1][a] * qY;   // (R^T)[a][0..1]
                    const size_t o = (i * nr + j) * nc + k;
                    md = std::fmax(md, std::fabs(g.x[o] - q[0]));
                    md = std::fmax(md, std::fabs(g.y[o] - q[1]));
                    md = std::fmax(md, std::fabs(g.z[o] - q[2]));
                    const bool out = std::fabs(q[0]) > M_PI || std::fabs(q[1]) > M_PI || std::fabs(q[2]) > M_PI;
                    masked_ref += out;
                    masked += g.w[o] == 0.0;
                }
        }
        std::printf("  PolarGrid max |point - R^T q| %.3e, masked %zu of %zu\n", md, masked, g.npts);
        CHECK("PolarGrid points and mask", g.npts == np * nr * nc && md < 1e-13 && masked == masked_ref &&
                                               masked > 0);
    }

    // ---- alignment override (include/config.h:284-343) ----
    {
        Dataset_t<double> ds;
        ds.angles = {0.0, 1.0, 2.0};
        ds.gamma = -0.5;
        ds.shifts = {{0.25, -0.5}};   // a cor-offset: broadcast
        AlignmentTable<double> t;
        t.corrected_angles_deg = {0.0, 120.0, 240.0};
        Dataset_t<double> a = ds;
        apply_alignment(a, t);
        bool ok = a.shifts.size() == 3 && a.shifts[2][0] == 0.25 && a.shifts[2][1] == -0.5 &&
                  a.gamma == -0.5 && std::fabs(a.angles[1] - 120.0 * M_PI / 180) < 1e-15;
        t.has_gamma = t.has_beta = t.has_shifts = true;
        t.gamma_deg = 29.5;
        t.beta_deg = -0.25;
        t.shifts_px = {{1, 2}, {3, 4}, {5, 6}};
        Dataset_t<double> b = ds;
        apply_alignment(b, t);
        ok = ok && b.gamma == -29.5 * M_PI / 180 && b.beta == -0.25 * M_PI / 180 && b.shifts[1][0] == 3;
        t.corrected_angles_deg = {0.0, 90.0};
        Dataset_t<double> c = ds;
        bool threw = false;
        try {
            apply_alignment(c, t);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        CHECK("apply_alignment", ok && threw);
    }

    std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}

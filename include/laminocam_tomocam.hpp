/*
 * laminocam_tomocam.hpp - the reference's C++ hot-path API re-implemented over
 * the C ABI of liblaminocam.so (laminocam.h), header only.
 *
 * Names and signatures follow the reference headers they replace:
 *   include/polar_grid2d.h:35-66     cpu::PolarGrid2D
 *   include/polar_grid.h:35-94       cpu::PolarGrid
 *   include/toeplitz2d.h:40-126      cpu::PointSpreadFunction2D
 *   include/toeplitz.h:35-130        cpu::PointSpreadFunction
 *   include/projection.h:38-68       forward, backproj (2 overloads), sysmat (2),
 *                                    forward2d, backproj2d, sysmat2d (2)
 *   include/optimize.h:31-91         opt::Function, opt::cgsolver, opt::split_bregman
 *   include/bregman.h:34-55          opt::grad_u, opt::divergence, opt::laplacian
 *   include/mbir2d_prep.h:30-65      Sinogram2D, prepare_sinogram
 *   include/tomocam.h:31-33          MBIR (src/mbir.cpp:34-161)
 *   include/gpu/tomocam.h:60-75      gpu::MBIR, gpu::MBIR2D (src/gpu/mbir2d.cu:36-85)
 *   include/gpu/tomocam.h:91-93      gpu::volume_fits_one_gpu
 *   include/config.h:284-343         broadcast_shifts, AlignmentTable, apply_alignment
 * Errors: a nonzero status becomes std::invalid_argument (bad arguments,
 * unsupported shapes) or std::runtime_error (device failures), as the
 * reference throws (src/projection.cpp:105-108, include/finufft_plan.h:106).
 *
 * Array types.  The functions use only dims() / data() / size() and the
 * Array(dims_t) constructor of tomocam::Array<T> (include/array.h:61-110), so
 * they compile against the reference's array.h when it is included first.
 * Defining LAMINOCAM_STANDALONE_ARRAY instead provides a minimal
 * tomocam::dims_t / tomocam::Array<T> with that same interface, and the
 * driver types Dataset_t, Regularizer, ReconParams and OutputParams of
 * include/recon_params.h:32-93 (otherwise include that header first).
 *
 * Precision: Array<float> runs the FP32 path (FP64 reductions, PSF and
 * twiddles).  Array<double> runs the FP64 path of the Toeplitz operator,
 * kernel_hat, the solvers and the finite differences (the _f64 entry points),
 * as the reference's double instantiations do; the NUFFT projectors compute in
 * FP32 for either type, double arrays being converted at the boundary.
 *
 * Solvers: opt::Function<T> is the reference's std::function alias.  cgsolver
 * and split_bregman run the whole iteration on the device when the function
 * object holds an opt::Toeplitz2D / opt::Toeplitz3D (made by
 * opt::toeplitz(psf)); any other callable, such as the reference's own
 * [&psf](x) { return sysmat2d(x, psf); }, is applied as given by a host loop
 * that follows src/conjgrad.cpp:33-76 and src/bregman.cpp:36-94, with FP64
 * inner products and norms.
 */
#ifndef LAMINOCAM_TOMOCAM_HPP
#define LAMINOCAM_TOMOCAM_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "laminocam.h"

#ifdef LAMINOCAM_STANDALONE_ARRAY
namespace tomocam {
struct dims_t {
    size_t n1 = 0, n2 = 0, n3 = 0;
    dims_t() = default;
    dims_t(size_t a, size_t b, size_t c) : n1(a), n2(b), n3(c) {}
    size_t size() const { return n1 * n2 * n3; }
    bool operator==(const dims_t &o) const { return n1 == o.n1 && n2 == o.n2 && n3 == o.n3; }
};

template <typename T> class Array {
  public:
    Array() = default;
    explicit Array(dims_t d) : dims_(d), v_(d.size()) {}
    static Array zeros(dims_t d) { return Array(d); }
    dims_t dims() const { return dims_; }
    size_t size() const { return v_.size(); }
    size_t nslices() const { return dims_.n1; }
    size_t nrows() const { return dims_.n2; }
    size_t ncols() const { return dims_.n3; }
    T *data() { return v_.data(); }
    const T *data() const { return v_.data(); }
    T *begin() { return v_.data(); }
    const T *begin() const { return v_.data(); }
    T *end() { return v_.data() + v_.size(); }
    const T *end() const { return v_.data() + v_.size(); }
    T &operator[](size_t i) { return v_[i]; }
    const T &operator[](size_t i) const { return v_[i]; }

    // the element-wise arithmetic of include/array.h:146-235
    Array clone() const { return *this; }
    static Array ones(dims_t d) {
        Array a(d);
        for (auto &v : a.v_) v = T(1);
        return a;
    }
    Array &operator*=(T s) {
        for (auto &v : v_) v *= s;
        return *this;
    }
    Array operator*(T s) const { return Array(*this) *= s; }
    Array &operator/=(T s) {
        if (s == T(0)) throw std::runtime_error("Division by zero in Array<T>::operator/=");
        for (auto &v : v_) v /= s;
        return *this;
    }
    Array operator/(T s) const { return Array(*this) /= s; }
    Array &operator+=(const Array &o) {
        check_(o);
        for (size_t i = 0; i < v_.size(); ++i) v_[i] += o.v_[i];
        return *this;
    }
    Array operator+(const Array &o) const { return Array(*this) += o; }
    Array &operator-=(const Array &o) {
        check_(o);
        for (size_t i = 0; i < v_.size(); ++i) v_[i] -= o.v_[i];
        return *this;
    }
    Array operator-(const Array &o) const { return Array(*this) -= o; }
    Array &operator*=(const Array &o) {
        check_(o);
        for (size_t i = 0; i < v_.size(); ++i) v_[i] *= o.v_[i];
        return *this;
    }
    Array operator*(const Array &o) const { return Array(*this) *= o; }

  private:
    void check_(const Array &o) const {
        if (!(o.dims_ == dims_)) throw std::invalid_argument("Array: dimensions differ");
    }
    dims_t dims_;
    std::vector<T> v_;
};

// include/array_ops.h:63-82: reductions (accumulated in FP64 here)
namespace array {
template <typename T> T dot(const Array<T> &a, const Array<T> &b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += (double)a[i] * (double)b[i];
    return (T)s;
}
template <typename T> T norm2(const Array<T> &a) { return (T)std::sqrt((double)dot(a, a)); }
} // namespace array
} // namespace tomocam
#endif

namespace laminocam {

inline void check(int status) {
    if (status == LC_OK) return;
    const std::string msg = lc_last_error();
    if (status == LC_EINVAL || status == LC_EUNSUP) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// one device context per process (device 0), created on first use
inline lc_ctx context() {
    static lc_ctx ctx = [] {
        lc_ctx c = nullptr;
        check(lc_ctx_create(0, &c));
        return c;
    }();
    return ctx;
}

// FP32 staging of a T array (no copy for float)
template <typename T> struct F32In {
    std::vector<float> buf;
    const float *p;
    F32In(const T *src, size_t n) {
        if constexpr (std::is_same<T, float>::value) {
            p = src;
        } else {
            buf.assign(src, src + n);
            p = buf.data();
        }
    }
};

template <typename T> struct F32Out {
    std::vector<float> buf;
    T *dst;
    float *p;
    size_t n;
    F32Out(T *d, size_t n_) : dst(d), n(n_) {
        if constexpr (std::is_same<T, float>::value) {
            p = d;
        } else {
            buf.resize(n);
            p = buf.data();
        }
    }
    void commit() {
        if constexpr (!std::is_same<T, float>::value)
            for (size_t i = 0; i < n; ++i) dst[i] = (T)buf[i];
    }
};

// element-type dispatch to the FP32 or FP64 entry points
inline int psf2d_kernel_hat(lc_ctx c, lc_geom2d g, float *o) { return lc_psf2d_kernel_hat(c, g, o); }
inline int psf2d_kernel_hat(lc_ctx c, lc_geom2d g, double *o) { return lc_psf2d_kernel_hat_f64(c, g, o); }
inline int psf3d_kernel_hat(lc_ctx c, lc_geom3d g, float *o) { return lc_psf3d_kernel_hat(c, g, o); }
inline int psf3d_kernel_hat(lc_ctx c, lc_geom3d g, double *o) { return lc_psf3d_kernel_hat_f64(c, g, o); }
inline int sysmat2d_psf(lc_ctx c, lc_geom2d g, const float *x, int64_t ns, float *y) {
    return lc_sysmat2d_psf(c, g, x, ns, y, 0);
}
inline int sysmat2d_psf(lc_ctx c, lc_geom2d g, const double *x, int64_t ns, double *y) {
    return lc_sysmat2d_psf_f64(c, g, x, ns, y, 0);
}
inline int sysmat3d_psf(lc_ctx c, lc_geom3d g, const float *x, float *y) { return lc_sysmat3d_psf(c, g, x, y, 0); }
inline int sysmat3d_psf(lc_ctx c, lc_geom3d g, const double *x, double *y) {
    return lc_sysmat3d_psf_f64(c, g, x, y, 0);
}
inline int cg2d(lc_ctx c, lc_geom2d g, const float *y, float *x, int64_t ns, int64_t it, double tol,
                double xtol, lc_report *r) {
    return lc_cg2d(c, g, y, x, ns, it, tol, xtol, 0, r);
}
inline int cg2d(lc_ctx c, lc_geom2d g, const double *y, double *x, int64_t ns, int64_t it, double tol,
                double xtol, lc_report *r) {
    return lc_cg2d_f64(c, g, y, x, ns, it, tol, xtol, 0, r);
}
inline int cg3d(lc_ctx c, lc_geom3d g, const float *y, float *x, int64_t it, double tol, double xtol,
                lc_report *r) {
    return lc_cg3d(c, g, y, x, it, tol, xtol, 0, r);
}
inline int cg3d(lc_ctx c, lc_geom3d g, const double *y, double *x, int64_t it, double tol, double xtol,
                lc_report *r) {
    return lc_cg3d_f64(c, g, y, x, it, tol, xtol, 0, r);
}
inline int sb2d(lc_ctx c, lc_geom2d g, const float *y, float *x, int64_t ns, double la, double mu,
                int64_t outer, int64_t inner, double tol, double xtol, lc_report *r) {
    return lc_split_bregman2d(c, g, y, x, ns, la, mu, outer, inner, tol, xtol, 0, r);
}
inline int sb2d(lc_ctx c, lc_geom2d g, const double *y, double *x, int64_t ns, double la, double mu,
                int64_t outer, int64_t inner, double tol, double xtol, lc_report *r) {
    return lc_split_bregman2d_f64(c, g, y, x, ns, la, mu, outer, inner, tol, xtol, 0, r);
}
inline int sb3d(lc_ctx c, lc_geom3d g, const float *y, float *x, double la, double mu, int64_t outer,
                int64_t inner, double tol, double xtol, lc_report *r) {
    return lc_split_bregman3d(c, g, y, x, la, mu, outer, inner, tol, xtol, 0, r);
}
inline int sb3d(lc_ctx c, lc_geom3d g, const double *y, double *x, double la, double mu, int64_t outer,
                int64_t inner, double tol, double xtol, lc_report *r) {
    return lc_split_bregman3d_f64(c, g, y, x, la, mu, outer, inner, tol, xtol, 0, r);
}
inline int fd(int op, lc_ctx c, const float *in, int64_t n1, int64_t n2, int64_t n3, float *out) {
    return op == 0 ? lc_grad(c, in, n1, n2, n3, out, 0)
                   : op == 1 ? lc_divergence(c, in, n1, n2, n3, out, 0) : lc_laplacian(c, in, n1, n2, n3, out, 0);
}
inline int fd(int op, lc_ctx c, const double *in, int64_t n1, int64_t n2, int64_t n3, double *out) {
    return op == 0 ? lc_grad_f64(c, in, n1, n2, n3, out, 0)
                   : op == 1 ? lc_divergence_f64(c, in, n1, n2, n3, out, 0)
                             : lc_laplacian_f64(c, in, n1, n2, n3, out, 0);
}

struct Geom2 {
    lc_geom2d h = nullptr;
    ~Geom2() { if (h) lc_geom2d_destroy(h); }
};
struct Geom3 {
    lc_geom3d h = nullptr;
    ~Geom3() { if (h) lc_geom3d_destroy(h); }
};

} // namespace laminocam

namespace tomocam::cpu {

// include/polar_grid2d.h:35-66.  The device generates the points
// x = r cos(theta), y = r sin(theta), r = (k + 1/2) 2 pi / ncols - pi, in its
// kernels from the angle table; the members npts, x, y hold the same points on
// the host, computed in T, for callers that read them.
template <typename T> struct PolarGrid2D {
    size_t npts = 0;
    Array<T> x;
    Array<T> y;
    std::vector<double> theta;
    size_t ncols_ = 0;
    size_t nangles() const { return theta.size(); }
    size_t nradial() const { return ncols_; }
    dims_t dims() const { return dims_t{1, theta.size(), ncols_}; }
    size_t size() const { return theta.size() * ncols_; }
    PolarGrid2D() = default;
    PolarGrid2D(const std::vector<T> &th, size_t ncols)
        : x(dims_t{1, th.size(), ncols}), y(dims_t{1, th.size(), ncols}), theta(th.begin(), th.end()),
          ncols_(ncols) {
        npts = th.size() * ncols;
        const T dr = T(2 * M_PI) / static_cast<T>(ncols);
        for (size_t i = 0; i < th.size(); ++i) {
            const T c = std::cos(th[i]), s = std::sin(th[i]);
            for (size_t k = 0; k < ncols; ++k) {
                const T r = T(k + 0.5) * dr - T(M_PI);
                x.data()[i * ncols + k] = c * r;
                y.data()[i * ncols + k] = s * r;
            }
        }
    }

    // device geometry for N x N images (PSF built on demand)
    lc_geom2d geometry(size_t N) const {
        auto &g = cache_[N];
        if (!g) {
            g = std::make_shared<laminocam::Geom2>();
            laminocam::check(lc_geom2d_create(laminocam::context(), theta.data(), (int64_t)theta.size(),
                                              (int64_t)ncols_, (int64_t)N, LC_GEOM_LAZY_PSF, &g->h));
        }
        return g->h;
    }

  private:
    mutable std::map<size_t, std::shared_ptr<laminocam::Geom2>> cache_;
};

// include/polar_grid.h:35-94: points R_i (qX, qY, 0), R_i = RotationTranspose(theta, gamma, beta)
// (include/rotation.h:28-56), qX = (k + 1/2) 2 pi / ncols - pi, qY likewise over
// nrows; w = 0 where a coordinate exceeds pi in magnitude, else 1.  The device
// forms the two in-plane columns of R_i in FP64 and generates the points in
// its kernels; npts, x, y, z, w hold them on the host, computed in T.
template <typename T> struct PolarGrid {
    size_t npts = 0;
    Array<T> x;
    Array<T> y;
    Array<T> z;
    Array<T> w;
    std::vector<double> theta, gamma, beta;
    size_t nrows_ = 0, ncols_ = 0;
    dims_t dims() const { return dims_t{theta.size(), nrows_, ncols_}; }
    size_t size() const { return theta.size() * nrows_ * ncols_; }
    PolarGrid() = default;
    PolarGrid(const std::vector<T> &th, const std::vector<T> &ga, const std::vector<T> &be, size_t nrows,
              size_t ncols)
        : theta(th.begin(), th.end()), gamma(ga.begin(), ga.end()), beta(be.begin(), be.end()),
          nrows_(nrows), ncols_(ncols) {
        if (ga.size() != th.size() || be.size() != th.size())
            throw std::invalid_argument("PolarGrid: theta, gamma and beta sizes differ");
        const dims_t d{th.size(), nrows, ncols};
        x = Array<T>(d);
        y = Array<T>(d);
        z = Array<T>(d);
        w = Array<T>(d);
        npts = d.size();
        const T dX = T(2 * M_PI) / static_cast<T>(ncols), dY = T(2 * M_PI) / static_cast<T>(nrows);
        for (size_t i = 0; i < th.size(); ++i) {
            const T ct = std::cos(th[i]), st = std::sin(th[i]);
            const T cb = std::cos(be[i]), sb = std::sin(be[i]);
            const T cg = std::cos(ga[i]), sg = std::sin(ga[i]);
            // columns 0 and 1 of the transposed rotation; q has no third component
            const T c0[3] = {cb * cg, -cb * sg, sb};
            const T c1[3] = {ct * sg + st * sb * cg, ct * cg - st * sb * sg, -st * cb};
            for (size_t j = 0; j < nrows; ++j) {
                const T qY = T(j + 0.5) * dY - T(M_PI);
                for (size_t k = 0; k < ncols; ++k) {
                    const T qX = T(k + 0.5) * dX - T(M_PI);
                    const size_t o = (i * nrows + j) * ncols + k;
                    const T px = c0[0] * qX + c1[0] * qY, py = c0[1] * qX + c1[1] * qY,
                            pz = c0[2] * qX + c1[2] * qY;
                    x.data()[o] = px;
                    y.data()[o] = py;
                    z.data()[o] = pz;
                    const T pi = T(M_PI);
                    w.data()[o] = (std::abs(px) > pi || std::abs(py) > pi || std::abs(pz) > pi) ? T(0) : T(1);
                }
            }
        }
    }

    lc_geom3d geometry(dims_t rd, bool lazy = true) const {
        const std::array<size_t, 3> key{rd.n1, rd.n2, rd.n3};
        auto &g = cache_[key];
        if (!g) {
            g = std::make_shared<laminocam::Geom3>();
            const int64_t d[3] = {(int64_t)rd.n1, (int64_t)rd.n2, (int64_t)rd.n3};
            laminocam::check(lc_geom3d_create(laminocam::context(), theta.data(), gamma.data(), beta.data(),
                                              (int64_t)theta.size(), (int64_t)nrows_, (int64_t)ncols_, d,
                                              lazy ? LC_GEOM_LAZY_PSF : 0, &g->h));
        }
        return g->h;
    }

  private:
    mutable std::map<std::array<size_t, 3>, std::shared_ptr<laminocam::Geom3>> cache_;
};

// include/toeplitz2d.h:40-126
template <typename T> class PointSpreadFunction2D {
  public:
    using complex_t = std::complex<T>;
    PointSpreadFunction2D() = default;
    PointSpreadFunction2D(const PolarGrid2D<T> &grid, size_t image_size) : N_(image_size) {
        g_ = std::make_shared<laminocam::Geom2>();
        laminocam::check(lc_geom2d_create(laminocam::context(), grid.theta.data(), (int64_t)grid.theta.size(),
                                          (int64_t)grid.nradial(), (int64_t)image_size, 0, &g_->h));
        int64_t info[5];
        laminocam::check(lc_geom2d_info(g_->h, info));
        M_ = (size_t)info[3];
    }
    size_t psf_size() const { return M_; }
    size_t image_size() const { return N_; }
    lc_geom2d handle() const { return g_->h; }

    Array<complex_t> kernel_hat() const {
        const size_t H = M_ / 2 + 1;
        std::vector<T> buf(2 * M_ * H);
        laminocam::check(laminocam::psf2d_kernel_hat(laminocam::context(), g_->h, buf.data()));
        Array<complex_t> out(dims_t{1, M_, H});
        for (size_t i = 0; i < M_ * H; ++i) out.data()[i] = complex_t((T)buf[2 * i], (T)buf[2 * i + 1]);
        return out;
    }

    Array<T> convolve(const Array<T> &x) const {
        const dims_t d = x.dims();
        if (d.n2 != N_ || d.n3 != N_) throw std::invalid_argument("convolve: slice size differs from the PSF's");
        Array<T> y(d);
        laminocam::check(laminocam::sysmat2d_psf(laminocam::context(), g_->h, x.data(), (int64_t)d.n1, y.data()));
        return y;
    }

  private:
    size_t N_ = 0, M_ = 0;
    std::shared_ptr<laminocam::Geom2> g_;
};

// include/toeplitz.h:35-130
template <typename T> class PointSpreadFunction {
  public:
    using complex_t = std::complex<T>;
    PointSpreadFunction() = default;
    PointSpreadFunction(const PolarGrid<T> &grid, dims_t recon_dims) : rd_(recon_dims) {
        g_ = std::make_shared<laminocam::Geom3>();
        const int64_t d[3] = {(int64_t)recon_dims.n1, (int64_t)recon_dims.n2, (int64_t)recon_dims.n3};
        laminocam::check(lc_geom3d_create(laminocam::context(), grid.theta.data(), grid.gamma.data(),
                                          grid.beta.data(), (int64_t)grid.theta.size(), (int64_t)grid.dims().n2,
                                          (int64_t)grid.dims().n3, d, 0, &g_->h));
        int64_t info[9];
        laminocam::check(lc_geom3d_info(g_->h, info));
        dims_ = dims_t{(size_t)info[6], (size_t)info[7], (size_t)info[8]};
    }
    dims_t dims() const { return dims_; }
    lc_geom3d handle() const { return g_->h; }

    Array<complex_t> kernel_hat() const {
        const size_t H = dims_.n3 / 2 + 1, n = dims_.n1 * dims_.n2 * H;
        std::vector<T> buf(2 * n);
        laminocam::check(laminocam::psf3d_kernel_hat(laminocam::context(), g_->h, buf.data()));
        Array<complex_t> out(dims_t{dims_.n1, dims_.n2, H});
        for (size_t i = 0; i < n; ++i) out.data()[i] = complex_t((T)buf[2 * i], (T)buf[2 * i + 1]);
        return out;
    }

    Array<T> convolve(const Array<T> &x) const {
        if (!(x.dims() == rd_)) throw std::invalid_argument("convolve: volume dims differ from the PSF's");
        Array<T> y(x.dims());
        laminocam::check(laminocam::sysmat3d_psf(laminocam::context(), g_->h, x.data(), y.data()));
        return y;
    }

  private:
    dims_t rd_, dims_;
    std::shared_ptr<laminocam::Geom3> g_;
};

} // namespace tomocam::cpu

namespace tomocam {

// include/projection.h:56-58, src/projection.cpp:146-160
template <typename T> Array<T> forward2d(const Array<T> &volume, const cpu::PolarGrid2D<T> &grid) {
    const dims_t d = volume.dims();
    if (d.n2 != d.n3) throw std::invalid_argument("forward2d: slices must be square");
    Array<T> out(dims_t{d.n1, grid.nangles(), grid.nradial()});
    laminocam::F32In<T> in(volume.data(), volume.size());
    laminocam::F32Out<T> o(out.data(), out.size());
    laminocam::check(lc_forward2d(laminocam::context(), grid.geometry(d.n2), in.p, (int64_t)d.n1, o.p, 0));
    o.commit();
    return out;
}

// include/projection.h:59-61, src/projection.cpp:168-182
template <typename T>
Array<T> backproj2d(const Array<T> &sino, const cpu::PolarGrid2D<T> &grid, const dims_t &recon_dims) {
    const dims_t d = sino.dims();
    if (d.n2 != grid.nangles() || d.n3 != grid.nradial())
        throw std::invalid_argument("backproj2d: sinogram does not match the grid");
    if (recon_dims.n1 != d.n1 || recon_dims.n2 != recon_dims.n3)
        throw std::invalid_argument("backproj2d: recon dims must be {nslices, N, N}");
    Array<T> out(recon_dims);
    laminocam::F32In<T> in(sino.data(), sino.size());
    laminocam::F32Out<T> o(out.data(), out.size());
    laminocam::check(
        lc_backproj2d(laminocam::context(), grid.geometry(recon_dims.n2), in.p, (int64_t)d.n1, o.p, 0));
    o.commit();
    return out;
}

// src/gradient.cpp:59-69 (NUFFT form) and :76-79 (Toeplitz form)
template <typename T> Array<T> sysmat2d(const Array<T> &x, const cpu::PolarGrid2D<T> &grid) {
    const dims_t d = x.dims();
    if (d.n2 != d.n3) throw std::invalid_argument("sysmat2d: slices must be square");
    Array<T> out(d);
    laminocam::F32In<T> in(x.data(), x.size());
    laminocam::F32Out<T> o(out.data(), out.size());
    laminocam::check(lc_sysmat2d_nufft(laminocam::context(), grid.geometry(d.n2), in.p, (int64_t)d.n1, o.p, 0));
    o.commit();
    return out;
}

template <typename T> Array<T> sysmat2d(const Array<T> &x, const cpu::PointSpreadFunction2D<T> &psf) {
    return psf.convolve(x);
}

// include/projection.h:38-39, src/projection.cpp:42-64
template <typename T> Array<T> forward(const Array<T> &volume, const cpu::PolarGrid<T> &grid) {
    Array<T> out(grid.dims());
    laminocam::F32In<T> in(volume.data(), volume.size());
    laminocam::F32Out<T> o(out.data(), out.size());
    laminocam::check(lc_forward3d(laminocam::context(), grid.geometry(volume.dims()), in.p, o.p, 0));
    o.commit();
    return out;
}

// include/projection.h:41-47, src/projection.cpp:71-131
template <typename T>
Array<T> backproj(const Array<T> &proj, const cpu::PolarGrid<T> &grid, const dims_t &recon_dims) {
    if (!(proj.dims() == grid.dims())) throw std::invalid_argument("backproj: projections do not match the grid");
    Array<T> out(recon_dims);
    laminocam::F32In<T> in(proj.data(), proj.size());
    laminocam::F32Out<T> o(out.data(), out.size());
    laminocam::check(lc_backproj3d(laminocam::context(), grid.geometry(recon_dims), in.p, nullptr, o.p, 0));
    o.commit();
    return out;
}

template <typename T>
Array<T> backproj(const Array<T> &proj, const cpu::PolarGrid<T> &grid, const dims_t &recon_dims,
                  const std::vector<std::array<T, 2>> &shifts) {
    if (shifts.size() != proj.nslices())
        throw std::invalid_argument("shifts size must match the number of slices in the projection");
    if (!(proj.dims() == grid.dims())) throw std::invalid_argument("backproj: projections do not match the grid");
    std::vector<double> sh(2 * shifts.size());
    for (size_t i = 0; i < shifts.size(); ++i) {
        sh[2 * i] = (double)shifts[i][0];
        sh[2 * i + 1] = (double)shifts[i][1];
    }
    Array<T> out(recon_dims);
    laminocam::F32In<T> in(proj.data(), proj.size());
    laminocam::F32Out<T> o(out.data(), out.size());
    laminocam::check(lc_backproj3d(laminocam::context(), grid.geometry(recon_dims), in.p, sh.data(), o.p, 0));
    o.commit();
    return out;
}

// src/gradient.cpp:32-41 (NUFFT form, no mask) and :49-52 (Toeplitz form)
template <typename T> Array<T> sysmat(const Array<T> &x, const cpu::PolarGrid<T> &grid) {
    Array<T> out(x.dims());
    laminocam::F32In<T> in(x.data(), x.size());
    laminocam::F32Out<T> o(out.data(), out.size());
    laminocam::check(lc_sysmat3d_nufft(laminocam::context(), grid.geometry(x.dims()), in.p, o.p, 0));
    o.commit();
    return out;
}

template <typename T> Array<T> sysmat(const Array<T> &x, const cpu::PointSpreadFunction<T> &psf) {
    return psf.convolve(x);
}

} // namespace tomocam

namespace tomocam::opt {

// include/optimize.h:31-61
template <typename T> using Function = std::function<Array<T>(const Array<T> &)>;
template <typename T> using Residual = std::function<T(const Array<T> &)>;

enum class Optimizer { SPLIT_BREGMAN, CG_SOLVER };

template <typename T> struct OptimizerConfig {
    Optimizer method = Optimizer::SPLIT_BREGMAN;
    T lambda = T(0.1);
    T mu = T(5.0);
    T tol = T(1e-5);
    T xtol = T(1e-5);
    size_t outer_max = 50;
    size_t inner_max = 4;
    OptimizerConfig() = default;
    OptimizerConfig(Optimizer meth, T lam, T m, size_t out_max, size_t in_max, T tol1, T tol2)
        : method(meth), lambda(lam), mu(m), tol(tol1), xtol(tol2), outer_max(out_max), inner_max(in_max) {}
};

// The Toeplitz normal operator x -> sysmat(x, psf) as a function object.  A
// Function<T> holding one of these is recognised by cgsolver and
// split_bregman (std::function::target), which then keep the iteration on the
// device.
template <typename T> struct Toeplitz2D {
    cpu::PointSpreadFunction2D<T> psf;
    Array<T> operator()(const Array<T> &x) const { return psf.convolve(x); }
};
template <typename T> struct Toeplitz3D {
    cpu::PointSpreadFunction<T> psf;
    Array<T> operator()(const Array<T> &x) const { return psf.convolve(x); }
};
template <typename T> Toeplitz2D<T> toeplitz(const cpu::PointSpreadFunction2D<T> &psf) { return {psf}; }
template <typename T> Toeplitz3D<T> toeplitz(const cpu::PointSpreadFunction<T> &psf) { return {psf}; }

// include/bregman.h:34-55, src/finitediff.cpp:26-90 (interior-only differences)
template <typename T> std::array<Array<T>, 3> grad_u(const Array<T> &u) {
    const dims_t d = u.dims();
    std::vector<T> g(3 * u.size());
    laminocam::check(laminocam::fd(0, laminocam::context(), u.data(), (int64_t)d.n1, (int64_t)d.n2,
                                   (int64_t)d.n3, g.data()));
    std::array<Array<T>, 3> out{Array<T>(d), Array<T>(d), Array<T>(d)};
    for (int c = 0; c < 3; ++c) std::memcpy(out[c].data(), g.data() + c * u.size(), sizeof(T) * u.size());
    return out;
}

template <typename T> Array<T> divergence(const std::array<Array<T>, 3> &g) {
    const dims_t d = g[0].dims();
    const size_t n = g[0].size();
    std::vector<T> in(3 * n);
    for (int c = 0; c < 3; ++c) std::memcpy(in.data() + c * n, g[c].data(), sizeof(T) * n);
    Array<T> out(d);
    laminocam::check(laminocam::fd(1, laminocam::context(), in.data(), (int64_t)d.n1, (int64_t)d.n2,
                                   (int64_t)d.n3, out.data()));
    return out;
}

template <typename T> Array<T> laplacian(const Array<T> &u) {
    const dims_t d = u.dims();
    Array<T> out(d);
    laminocam::check(laminocam::fd(2, laminocam::context(), u.data(), (int64_t)d.n1, (int64_t)d.n2,
                                   (int64_t)d.n3, out.data()));
    return out;
}
namespace detail {
template <typename T> double dot(const T *a, const T *b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return s;
}

template <typename T> Array<T> call_op(const Function<T> &A, const Array<T> &x) {
    if (!A) throw std::invalid_argument("the operator is an empty function");
    Array<T> y = A(x);
    if (y.size() != x.size()) throw std::invalid_argument("the operator changed the array size");
    return y;
}

// src/conjgrad.cpp:33-76 on the host, for an operator given only as a
// callable: r = y - A x, p = r; per iteration alpha = <r,r>/<p,Ap> (stop when
// |<p,Ap>| < 1e-10), x += alpha p, r -= alpha Ap, p = r + (<r',r'>/<r,r>) p;
// stop when ||r||/(||y||+1e-8) < tol or ||alpha p||/(||x||+1e-8) < xtol.
// Inner products and norms accumulate in FP64.
template <typename T>
Array<T> cg_host(const Function<T> &A, const Array<T> &y, const Array<T> &x0, size_t max_iter, T tol, T xtol) {
    const size_t n = y.size();
    if (x0.size() != n) throw std::invalid_argument("cgsolver: x0 and y differ in size");
    Array<T> x(x0.dims());
    std::memcpy(x.data(), x0.data(), sizeof(T) * n);
    Array<T> r = call_op(A, x);
    for (size_t i = 0; i < n; ++i) r.data()[i] = y.data()[i] - r.data()[i];
    const double ynorm = std::sqrt(dot(y.data(), y.data(), n));
    Array<T> p(x0.dims());
    std::memcpy(p.data(), r.data(), sizeof(T) * n);
    double rs_old = dot(r.data(), r.data(), n);
    for (size_t it = 0; it < max_iter; ++it) {
        const Array<T> Ap = call_op(A, p);
        const double pAp = dot(p.data(), Ap.data(), n);
        if (std::fabs(pAp) < 1.0e-10) break;
        const double alpha = rs_old / pAp;
        double step2 = 0.0;
        T *xp = x.data(), *rp = r.data(), *pp = p.data();
        const T *ap = Ap.data();
        for (size_t i = 0; i < n; ++i) {
            const T st = (T)(pp[i] * alpha);
            step2 += (double)st * st;
            xp[i] += st;
            rp[i] -= (T)(ap[i] * alpha);
        }
        const double rs_new = dot(rp, rp, n);
        const double beta = rs_new / rs_old;
        for (size_t i = 0; i < n; ++i) pp[i] = rp[i] + (T)(pp[i] * beta);
        rs_old = rs_new;
        const double res = std::sqrt(rs_new) / (ynorm + 1e-8);
        const double dx = std::sqrt(step2) / (std::sqrt(dot(xp, xp, n)) + 1e-8);
        if (res < (double)tol || dx < (double)xtol) break;
    }
    return x;
}

// src/bregman.cpp:36-94 on the host around an operator given only as a
// callable; the finite differences run on the device (grad_u, divergence,
// laplacian above).  Per outer iteration: rhs = y - mu div(d - b);
// x = CG(A - mu laplacian, rhs, x, inner_max); g = grad_u(x);
// sk = sqrt(sum_c (g_c + b_c)^2); d_c = max(0, sk - lambda/mu) (g_c + b_c) /
// (sk + 1e-8); b += g - d; stop when ||x - x_old|| / (||x_old|| + 1e-8) < xtol.
template <typename T>
Array<T> sb_host(const Function<T> &A, const Array<T> &y, const Array<T> &x0, T lambda, T mu, size_t outer_max,
                 size_t inner_max, T tol, T xtol) {
    const dims_t dm = x0.dims();
    const size_t n = x0.size();
    if (y.size() != n) throw std::invalid_argument("split_bregman: x0 and y differ in size");
    Array<T> x(dm), x_old(dm);
    std::memcpy(x.data(), x0.data(), sizeof(T) * n);
    std::memcpy(x_old.data(), x0.data(), sizeof(T) * n);
    std::array<Array<T>, 3> d{Array<T>(dm), Array<T>(dm), Array<T>(dm)};
    std::array<Array<T>, 3> b{Array<T>(dm), Array<T>(dm), Array<T>(dm)};
    for (int c = 0; c < 3; ++c) {
        std::fill(d[c].data(), d[c].data() + n, T(0));
        std::fill(b[c].data(), b[c].data() + n, T(0));
    }
    const Function<T> Ap = [&A, mu](const Array<T> &u) {
        Array<T> out = call_op(A, u);
        const Array<T> lap = laplacian(u);
        for (size_t i = 0; i < out.size(); ++i) out.data()[i] -= lap.data()[i] * mu;
        return out;
    };
    const T thr = lambda / mu;
    for (size_t it = 0; it < outer_max; ++it) {
        std::array<Array<T>, 3> dmb{Array<T>(dm), Array<T>(dm), Array<T>(dm)};
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < n; ++i) dmb[c].data()[i] = d[c].data()[i] - b[c].data()[i];
        const Array<T> dv = divergence(dmb);
        Array<T> rhs(dm);
        for (size_t i = 0; i < n; ++i) rhs.data()[i] = y.data()[i] - dv.data()[i] * mu;
        Array<T> xn = cg_host(Ap, rhs, x, inner_max, tol, xtol);
        std::memcpy(x.data(), xn.data(), sizeof(T) * n);
        const std::array<Array<T>, 3> g = grad_u(x);
        for (size_t i = 0; i < n; ++i) {
            T v[3];
            double ss = 0.0;
            for (int c = 0; c < 3; ++c) {
                v[c] = g[c].data()[i] + b[c].data()[i];
                ss += (double)v[c] * v[c];
            }
            const T sk = (T)std::sqrt(ss);
            const T sh = sk - thr > T(0) ? sk - thr : T(0);
            for (int c = 0; c < 3; ++c) d[c].data()[i] = sh * (v[c] / (sk + T(1e-8)));
        }
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < n; ++i) b[c].data()[i] += g[c].data()[i] - d[c].data()[i];
        double dn = 0.0, on = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double df = (double)x.data()[i] - (double)x_old.data()[i];
            dn += df * df;
            on += (double)x_old.data()[i] * x_old.data()[i];
        }
        std::memcpy(x_old.data(), x.data(), sizeof(T) * n);
        if (std::sqrt(dn) / (std::sqrt(on) + 1e-8) < (double)xtol) break;
    }
    return x;
}
} // namespace detail

// The operator parameter is a non-deduced context: T comes from the arrays, so
// a lambda or an opt::toeplitz(psf) object converts to Function<T> in place.
// The function type is the reference's all the same.
template <typename T> struct same_type { using type = T; };
template <typename T> using OperatorArg = typename same_type<Function<T>>::type;

// include/optimize.h:89-91, src/conjgrad.cpp:33-76
template <typename T>
Array<T> cgsolver(const OperatorArg<T> &A, const Array<T> &y, const Array<T> &x, size_t max_iter, T tol, T xtol) {
    const auto *t2 = A.template target<Toeplitz2D<T>>();
    const auto *t3 = A.template target<Toeplitz3D<T>>();
    if (!t2 && !t3) return detail::cg_host(A, y, x, max_iter, tol, xtol);
    if (x.size() != y.size()) throw std::invalid_argument("cgsolver: x0 and y differ in size");
    Array<T> out(x.dims());
    std::memcpy(out.data(), x.data(), sizeof(T) * x.size());
    lc_report rep;
    if (t2)
        laminocam::check(laminocam::cg2d(laminocam::context(), t2->psf.handle(), y.data(), out.data(),
                                         (int64_t)y.dims().n1, (int64_t)max_iter, (double)tol, (double)xtol,
                                         &rep));
    else
        laminocam::check(laminocam::cg3d(laminocam::context(), t3->psf.handle(), y.data(), out.data(),
                                         (int64_t)max_iter, (double)tol, (double)xtol, &rep));
    return out;
}

// include/optimize.h:74-77, src/bregman.cpp:36-94 (isotropic shrink,
// interior-only differences as the reference CPU path)
template <typename T>
Array<T> split_bregman(const OperatorArg<T> &A, const Array<T> &y, const Array<T> &x0, T lambda, T mu,
                       size_t outer_max, size_t inner_max, T tol, T xtol) {
    const auto *t2 = A.template target<Toeplitz2D<T>>();
    const auto *t3 = A.template target<Toeplitz3D<T>>();
    if (!t2 && !t3) return detail::sb_host(A, y, x0, lambda, mu, outer_max, inner_max, tol, xtol);
    if (x0.size() != y.size()) throw std::invalid_argument("split_bregman: x0 and y differ in size");
    Array<T> out(x0.dims());
    std::memcpy(out.data(), x0.data(), sizeof(T) * x0.size());
    lc_report rep;
    if (t2)
        laminocam::check(laminocam::sb2d(laminocam::context(), t2->psf.handle(), y.data(), out.data(),
                                         (int64_t)y.dims().n1, (double)lambda, (double)mu, (int64_t)outer_max,
                                         (int64_t)inner_max, (double)tol, (double)xtol, &rep));
    else
        laminocam::check(laminocam::sb3d(laminocam::context(), t3->psf.handle(), y.data(), out.data(),
                                         (double)lambda, (double)mu, (int64_t)outer_max, (int64_t)inner_max,
                                         (double)tol, (double)xtol, &rep));
    return out;
}

} // namespace tomocam::opt

// ---------------------------------------------------------------------------
// whole-solve drivers
// ---------------------------------------------------------------------------
#ifdef LAMINOCAM_STANDALONE_ARRAY
namespace tomocam {
// include/recon_params.h:32-54: the fields the drivers read
template <typename T> struct Dataset_t {
    Array<T> projs;                         // {n_angles, n_rows, n_cols}
    std::vector<T> angles;
    T gamma;
    T beta = T(0);
    std::vector<std::array<T, 2>> shifts;   // per-projection [dx, dy] in pixels
};
enum class Regularizer { SPLIT_BREGMAN, UNCONSTRAINED };
struct ReconParams {
    Regularizer regularizer = Regularizer::UNCONSTRAINED;
    std::array<size_t, 3> recon_dims = {0, 0, 0};
    size_t maxIters = 100;
    size_t innerIters = 1;
    float lambda = 0.1f;
    float mu = 10.0f;
    float tol = 1e-5f;
    float xtol = 1e-5f;
    float PAD_FACTOR = 1.4142f;
    // the reference's multi-device streaming fields (include/recon_params.h:56-60);
    // read from configuration files, this library runs one device
    size_t numGpus = 0;
    size_t chunkSlices = 0;
    bool streamed = false;

    // include/recon_params.h:62-81
    void print(std::ostream &os) const {
        const bool sb = regularizer == Regularizer::SPLIT_BREGMAN;
        os << "Reconstruction Parameters:\n";
        os << "  max_outer_iters: " << maxIters << "\n";
        os << "  recon_dims: [" << recon_dims[0] << ", " << recon_dims[1] << ", " << recon_dims[2] << "]\n";
        os << "  tol: " << tol << "\n";
        os << "  xtol: " << xtol << "\n";
        os << "  regularizer: " << (sb ? "Split-Bregman" : "Unconstrained") << "\n";
        if (sb) {
            os << "    inner_iters: " << innerIters << "\n";
            os << "    lambda: " << lambda << "\n";
            os << "    mu: " << mu << "\n";
        }
        if (numGpus > 0) os << "  num_gpus: " << numGpus << "\n";
        if (chunkSlices > 0) os << "  chunk_slices: " << chunkSlices << "\n";
        if (streamed) os << "  streamed: true\n";
    }
};

// include/recon_params.h:84-93
struct OutputParams {
    std::string filepath;
    std::vector<std::string> formats;
    bool has_format(const std::string &format) const {
        return std::find(formats.begin(), formats.end(), format) != formats.end();
    }
};
} // namespace tomocam
#endif

namespace laminocam {

inline lc_recon_params recon_params(const tomocam::ReconParams &p) {
    lc_recon_params c;
    lc_recon_params_default(&c);
    c.regularizer = p.regularizer == tomocam::Regularizer::SPLIT_BREGMAN ? LC_REG_SPLIT_BREGMAN
                                                                         : LC_REG_UNCONSTRAINED;
    for (int i = 0; i < 3; ++i) c.recon_dims[i] = (int64_t)p.recon_dims[i];
    c.max_iters = (int64_t)p.maxIters;
    c.inner_iters = (int64_t)p.innerIters;
    c.lambda = p.lambda;
    c.mu = p.mu;
    c.tol = p.tol;
    c.xtol = p.xtol;
    c.pad_factor = p.PAD_FACTOR;
    return c;
}

inline int prepare_sinogram(lc_ctx c, int32_t nd, const float *const *p, const int64_t *na, int64_t nr,
                            int64_t nc, float *out) {
    return lc_prepare_sinogram(c, nd, p, na, nr, nc, out, nullptr, 0);
}
inline int prepare_sinogram(lc_ctx c, int32_t nd, const double *const *p, const int64_t *na, int64_t nr,
                            int64_t nc, double *out) {
    return lc_prepare_sinogram_f64(c, nd, p, na, nr, nc, out, nullptr, 0);
}
inline int mbir2d(lc_ctx c, const float *s, int64_t ns, int64_t na, int64_t nc, const double *th,
                  const lc_recon_params *p, float *x) {
    return lc_mbir2d(c, s, ns, na, nc, th, p, x, nullptr, 0);
}
inline int mbir2d(lc_ctx c, const double *s, int64_t ns, int64_t na, int64_t nc, const double *th,
                  const lc_recon_params *p, double *x) {
    return lc_mbir2d_f64(c, s, ns, na, nc, th, p, x, nullptr, 0);
}
inline int mbir3d(lc_ctx c, int32_t nd, const float *const *pr, const int64_t *na, int64_t nr, int64_t nc,
                  const double *th, const double *ga, const double *be, const double *sh,
                  const lc_recon_params *p, float *x) {
    return lc_mbir3d(c, nd, pr, na, nr, nc, th, ga, be, sh, p, x, nullptr, nullptr, 0);
}
inline int mbir3d(lc_ctx c, int32_t nd, const double *const *pr, const int64_t *na, int64_t nr, int64_t nc,
                  const double *th, const double *ga, const double *be, const double *sh,
                  const lc_recon_params *p, double *x) {
    return lc_mbir3d_f64(c, nd, pr, na, nr, nc, th, ga, be, sh, p, x, nullptr, nullptr, 0);
}

// projection pointers and counts of the datasets; every dataset shares one detector size
template <typename T> struct DatasetView {
    std::vector<const T *> projs;
    std::vector<int64_t> nang;
    std::vector<double> theta;
    size_t nrows = 0, ncols = 0;
    explicit DatasetView(const std::vector<tomocam::Dataset_t<T>> &ds) {
        if (ds.empty()) throw std::invalid_argument("no datasets");
        nrows = ds[0].projs.nrows();
        ncols = ds[0].projs.ncols();
        for (const auto &d : ds) {
            if (d.projs.nrows() != nrows || d.projs.ncols() != ncols)
                throw std::invalid_argument("datasets differ in detector size");
            if (d.angles.size() != d.projs.nslices())
                throw std::invalid_argument("one angle per projection");
            projs.push_back(d.projs.data());
            nang.push_back((int64_t)d.projs.nslices());
            theta.insert(theta.end(), d.angles.begin(), d.angles.end());
        }
    }
};

} // namespace laminocam

namespace tomocam {

// include/config.h:284-296: the per-projection [dx, dy] shifts of one dataset;
// none gives zeros, one pair (the legacy 'cor-offset') is repeated for every
// projection, otherwise there is one pair per projection
template <typename T>
std::vector<std::array<T, 2>> broadcast_shifts(const std::vector<std::array<T, 2>> &shifts, size_t nproj) {
    if (shifts.empty()) return std::vector<std::array<T, 2>>(nproj, std::array<T, 2>{T(0), T(0)});
    if (shifts.size() == 1) return std::vector<std::array<T, 2>>(nproj, shifts[0]);
    if (shifts.size() != nproj)
        throw std::runtime_error("shifts file has " + std::to_string(shifts.size()) + " entries but " +
                                 std::to_string(nproj) + " projections were loaded");
    return shifts;
}

// The contents of an alignment file (include/config.h:298-343), parsed by
// the caller: optional tilts in degrees, the corrected angles in degrees, and
// optional per-projection shifts in pixels.
template <typename T> struct AlignmentTable {
    bool has_gamma = false, has_beta = false, has_shifts = false;
    T gamma_deg = T(0), beta_deg = T(0);
    std::vector<T> corrected_angles_deg;
    std::vector<std::array<T, 2>> shifts_px;
};

// include/config.h:298-343: the alignment override of a dataset.  gamma =
// -gamma_deg (the parser's sign convention) and beta = beta_deg, in radians,
// when given; the angles become the corrected angles; the shifts become
// shifts_px when given, else stay the dataset's (completed by
// broadcast_shifts), and must number one per corrected angle.
template <typename T> void apply_alignment(Dataset_t<T> &ds, const AlignmentTable<T> &a) {
    ds.shifts = broadcast_shifts<T>(ds.shifts, ds.angles.size());
    if (a.has_gamma) ds.gamma = -a.gamma_deg * T(M_PI) / T(180);
    if (a.has_beta) ds.beta = a.beta_deg * T(M_PI) / T(180);
    if (a.corrected_angles_deg.empty()) throw std::runtime_error("Alignment TOML missing 'corrected_angles_deg'");
    ds.angles.clear();
    for (T v : a.corrected_angles_deg) ds.angles.push_back(v * T(M_PI) / T(180));
    if (a.has_shifts) ds.shifts = a.shifts_px;
    if (ds.shifts.size() != ds.angles.size())
        throw std::runtime_error("Alignment TOML: shifts_px has " + std::to_string(ds.shifts.size()) +
                                 " entries but corrected_angles_deg has " + std::to_string(ds.angles.size()));
}

template <typename T> struct Sinogram2D {
    Array<T> data;          // {n_slices, n_angles, n_cols}
    std::vector<T> theta;   // one angle per projection (radians)
};

// include/mbir2d_prep.h:44-65: stack, divide by the global maximum (when > 0),
// {n_angles, n_rows, n_cols} -> {n_rows, n_angles, n_cols}; on the device
template <typename T> Sinogram2D<T> prepare_sinogram(const std::vector<Dataset_t<T>> &datasets) {
    laminocam::DatasetView<T> v(datasets);
    Sinogram2D<T> out;
    out.data = Array<T>(dims_t{v.nrows, v.theta.size(), v.ncols});
    laminocam::check(laminocam::prepare_sinogram(laminocam::context(), (int32_t)v.projs.size(),
                                                 v.projs.data(), v.nang.data(), (int64_t)v.nrows,
                                                 (int64_t)v.ncols, out.data.data()));
    out.theta.assign(v.theta.begin(), v.theta.end());
    return out;
}

// src/mbir.cpp:34-161: the 3-D laminography reconstruction, cropped to
// params.recon_dims.  As there, every projection needs a shift.
template <typename T> Array<T> MBIR(const std::vector<Dataset_t<T>> &datasets, const ReconParams &params) {
    laminocam::DatasetView<T> v(datasets);
    std::vector<double> gamma, beta, sh;
    for (const auto &d : datasets) {
        gamma.push_back((double)d.gamma);
        beta.push_back((double)d.beta);
        for (const auto &s : d.shifts) {
            sh.push_back((double)s[0]);
            sh.push_back((double)s[1]);
        }
    }
    if (sh.size() != 2 * v.theta.size())
        throw std::invalid_argument("shifts size must match the number of slices in the projection");
    const lc_recon_params p = laminocam::recon_params(params);
    Array<T> out(dims_t{params.recon_dims[0], params.recon_dims[1], params.recon_dims[2]});
    laminocam::check(laminocam::mbir3d(laminocam::context(), (int32_t)v.projs.size(), v.projs.data(),
                                       v.nang.data(), (int64_t)v.nrows, (int64_t)v.ncols, v.theta.data(),
                                       gamma.data(), beta.data(), sh.data(), &p, out.data()));
    return out;
}

namespace gpu {
// include/gpu/tomocam.h:91-93 (rule src/gpu/mbir2d_multi.cu:459-470): the
// single-device solvers' working set, Split-Bregman 27 N and CG 11 N elements
// of T, below 0.8 of the device's free memory
template <typename T> bool volume_fits_one_gpu(const ReconParams &params) {
    const lc_recon_params p = laminocam::recon_params(params);
    int32_t fits = 0;
    laminocam::check(lc_volume_fits_one_gpu(laminocam::context(), &p, (int32_t)sizeof(T), &fits));
    return fits != 0;
}

// include/gpu/tomocam.h:60-75: the device drivers are the same calls
template <typename T> Array<T> MBIR(const std::vector<Dataset_t<T>> &datasets, const ReconParams &params) {
    return tomocam::MBIR<T>(datasets, params);
}

// src/gpu/mbir2d.cu:36-85 (and cpu_mbir2d, src/mbir2d.cpp:66-100): the
// slice-stack reconstruction of a prepared sinogram {n_slices, n_angles, n_cols}
template <typename T>
Array<T> MBIR2D(const Array<T> &sinogram, const std::vector<T> &theta, const ReconParams &params) {
    if (theta.size() != sinogram.nrows()) throw std::invalid_argument("one angle per sinogram row");
    const std::vector<double> th(theta.begin(), theta.end());
    const lc_recon_params p = laminocam::recon_params(params);
    Array<T> out(dims_t{params.recon_dims[0], params.recon_dims[1], params.recon_dims[2]});
    laminocam::check(laminocam::mbir2d(laminocam::context(), sinogram.data(), (int64_t)sinogram.nslices(),
                                       (int64_t)sinogram.nrows(), (int64_t)sinogram.ncols(), th.data(), &p,
                                       out.data()));
    return out;
}
} // namespace gpu

} // namespace tomocam

#endif // LAMINOCAM_TOMOCAM_HPP

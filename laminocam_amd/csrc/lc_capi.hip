// extern "C" boundary of the library (include/laminocam.h).
#include <atomic>
#include <cstring>
#include <string>

#include "laminocam.h"
#include "lc_engine.hpp"
#include "lc_server.hpp"

using namespace lc;

std::string &lc::last_error() {
    thread_local std::string msg;
    return msg;
}

namespace {
template <typename F> int guard(F &&f) { return c_guard(std::forward<F>(f)); }
} // namespace

struct lc_ctx_s {
    Ctx c;
    // the caller's handle plus one share per geometry, batch and server
    // created on the context: it is freed when the last of them is destroyed
    std::atomic<int> refs{1};
    explicit lc_ctx_s(int dev) : c(dev) {}
};

static lc_ctx_s *ctx_retain(lc_ctx_s *c) {
    c->refs.fetch_add(1, std::memory_order_relaxed);
    return c;
}

static void ctx_release(lc_ctx_s *c) {
    if (c && c->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete c;
}

// waits for the object's context, destroys the object, then drops its share
// of the context
template <typename H> static void destroy_on_ctx(H *h) {
    if (!h) return;
    lc_ctx_s *c = h->ctx;
    c->c.set_current();
    c->c.sync();
    delete h;
    ctx_release(c);
}

struct lc_geom2d_s {
    lc_ctx_s *ctx;
    Geom2D g;
};

struct lc_geom3d_s {
    lc_ctx_s *ctx;
    Geom3D g;
};

struct lc_batch2d_s {
    lc_ctx_s *ctx;
    std::unique_ptr<Batch2D> b;
    std::vector<lc_report> reps;
};

// device address of buffer `which` (element type of the batch)
static void *batch_buf(Batch2D &b, int which) {
    switch (which) {
    case 0: return b.Xv();
    case 1: return b.Yv();
    case 2: return b.APv();
    default: throw Error(LC_EINVAL, "batch buffer index must be 0 (x), 1 (y) or 2 (Ax)");
    }
}

// element type T of the caller's arrays must match the batch's
template <typename T> static void batch_check_type(const Batch2D &b, const char *fn) {
    if (b.f64() != (sizeof(T) == 8))
        throw Error(LC_EINVAL, std::string(fn) + (b.f64() ? ": FP64 batch, use the _f64 entry point"
                                                           : ": FP32 batch, use the float entry point"));
}

static void fill_report(const CgState &s, lc_report *r) {
    r->iters = s.iters;
    r->outer = 0;
    r->res = s.res;
    r->dx = s.dx;
    r->stop = s.stop;
    r->pad = 0;
}

extern "C" {

int lc_abi_version(void) { return LC_ABI_VERSION; }

const char *lc_last_error(void) { return lc::last_error().c_str(); }

int lc_ctx_create(int device, lc_ctx *out) {
    return guard([&] {
        LC_REQUIRE(out, "lc_ctx_create: out is NULL");
        auto *c = new lc_ctx_s(device);   // (makes the device current)
        try {
            noalloc_setup(device);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

int lc_ctx_destroy(lc_ctx ctx) {
    return guard([&] { ctx_release(ctx); });
}

int lc_ctx_sync(lc_ctx ctx) {
    return guard([&] {
        LC_REQUIRE(ctx, "NULL context");
        ctx->c.set_current();
        ctx->c.sync();
    });
}

int lc_geom2d_create(lc_ctx ctx, const double *theta, int64_t nang, int64_t nrad,
                     int64_t image_n, uint32_t flags, lc_geom2d *out) {
    return guard([&] {
        LC_REQUIRE(ctx && theta && out, "lc_geom2d_create: NULL argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        auto *g = new lc_geom2d_s();
        g->ctx = ctx;
        try {
            geom2d_init(ctx->c, g->g, theta, nang, nrad, image_n);
            if (!(flags & LC_GEOM_LAZY_PSF)) geom2d_build_psf(ctx->c, g->g);
            ctx->c.sync();
        } catch (...) {
            delete g;
            throw;
        }
        ctx_retain(ctx);
        *out = g;
    });
}

int lc_geom2d_destroy(lc_geom2d g) {
    return guard([&] { destroy_on_ctx(g); });
}

int lc_geom2d_info(lc_geom2d g, int64_t *info) {
    return guard([&] {
        LC_REQUIRE(g && info, "lc_geom2d_info: NULL argument");
        info[0] = g->g.nang;
        info[1] = g->g.nrad;
        info[2] = g->g.N;
        info[3] = g->g.Mr;
        info[4] = g->g.Mr / 2 + 1;
    });
}

int lc_psf2d_kernel_hat(lc_ctx ctx, lc_geom2d g, float *out) {
    return guard([&] {
        LC_REQUIRE(ctx && g && out, "lc_psf2d_kernel_hat: NULL argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        geom2d_kernel_hat<float>(ctx->c, g->g, out);
    });
}

int lc_psf2d_kernel_hat_f64(lc_ctx ctx, lc_geom2d g, double *out) {
    return guard([&] {
        LC_REQUIRE(ctx && g && out, "lc_psf2d_kernel_hat_f64: NULL argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        geom2d_kernel_hat<double>(ctx->c, g->g, out);
    });
}

// one-problem solves and applies, shared by the 2-D and 3-D entry points;
// T = float runs the FP32 path, T = double the FP64 path
} // extern "C"

template <typename T>
static void sysmat_psf(lc_ctx ctx, ProbSpec ps, const T *x, T *y, uint32_t flags) {
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    ctx->c.set_current();
    Batch2D b(ctx->c, {ps}, sizeof(T) == 8);
    const size_t bytes = sizeof(T) * (size_t)b.size(0);   // exactly the problem's size
    const bool dev = flags & LC_DEVICE_PTRS;
    copy_in(ctx->c, b.Xv(), x, bytes, dev);
    b.toeplitz(b.Xv(), b.APv());
    copy_out(ctx->c, y, b.APv(), bytes, dev);
    ctx->c.sync();
}

extern "C" {

} // extern "C"

template <typename T>
static void cg_one(lc_ctx ctx, ProbSpec ps, const T *rhs, T *x, int64_t max_iter, double tol,
                   double xtol, uint32_t flags, lc_report *rep) {
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    ctx->c.set_current();
    Batch2D b(ctx->c, {ps}, sizeof(T) == 8);
    const size_t bytes = sizeof(T) * (size_t)b.size(0);
    const bool dev = flags & LC_DEVICE_PTRS;
    copy_in(ctx->c, b.Yv(), rhs, bytes, dev);
    if (!(flags & LC_X0_ZERO)) copy_in(ctx->c, b.Xv(), x, bytes, dev);
    const bool graph = flags & LC_GRAPH;
    b.cg({(int)max_iter}, tol, xtol, graph, !graph, (flags & LC_X0_ZERO) != 0);
    copy_out(ctx->c, x, b.Xv(), bytes, dev);
    ctx->c.sync();
    if (rep) fill_report(b.cg_state()[0], rep);
}

extern "C" {

} // extern "C"

template <typename T>
static void sb_one(lc_ctx ctx, ProbSpec ps, const T *rhs, T *x, double lambda, double mu,
                   int64_t outer, int64_t inner, double tol, double xtol, uint32_t flags,
                   lc_report *rep) {
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    ctx->c.set_current();
    Batch2D b(ctx->c, {ps}, sizeof(T) == 8);
    const size_t bytes = sizeof(T) * (size_t)b.size(0);
    const bool dev = flags & LC_DEVICE_PTRS;
    copy_in(ctx->c, b.Yv(), rhs, bytes, dev);
    copy_in(ctx->c, b.Xv(), x, bytes, dev);
    const bool graph = flags & LC_GRAPH;
    b.split_bregman({(float)lambda}, {(float)mu}, (int)outer, (int)inner, tol, xtol,
                    (flags & LC_TV_ZERO_PAD) ? 1 : 0, (flags & LC_TV_ANISO) != 0, graph, !graph);
    copy_out(ctx->c, x, b.Xv(), bytes, dev);
    ctx->c.sync();
    if (rep) {
        fill_report(b.cg_state()[0], rep);
        const SbState s = b.sb_state()[0];
        rep->outer = s.outer;
        rep->dx = s.norm_diff;
    }
}

extern "C" {

} // extern "C"

template <typename T>
static int sysmat2d_psf_t(lc_ctx ctx, lc_geom2d g, const T *x, int64_t nslices, T *y,
                          uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx && g && x && y, "lc_sysmat2d_psf: NULL argument");
        LC_REQUIRE(nslices >= 1, "lc_sysmat2d_psf: nslices must be >= 1");
        sysmat_psf<T>(ctx, ProbSpec{&g->g, (int)nslices}, x, y, flags);
    });
}

extern "C" {

int lc_sysmat2d_psf(lc_ctx ctx, lc_geom2d g, const float *x, int64_t nslices, float *y,
                    uint32_t flags) {
    return sysmat2d_psf_t(ctx, g, x, nslices, y, flags);
}

int lc_sysmat2d_psf_f64(lc_ctx ctx, lc_geom2d g, const double *x, int64_t nslices, double *y,
                        uint32_t flags) {
    return sysmat2d_psf_t(ctx, g, x, nslices, y, flags);
}

// forward2d (mode 0) / backproj2d (1) / sysmat2d NUFFT form (2); host arrays
// are staged through context scratch
static int project2d(lc_ctx ctx, lc_geom2d g, const float *in, int64_t nslices, float *out,
                     uint32_t flags, int mode) {
    return guard([&] {
        LC_REQUIRE(ctx && g && in && out, "projector: NULL argument");
        LC_REQUIRE(nslices >= 1 && nslices <= (1 << 20), "projector: bad nslices");
        LC_REQUIRE(g->ctx == ctx, "projector: geometry of another context");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        const Geom2D &gg = g->g;
        const int64_t nv = nslices * gg.N * gg.N, ns_ = nslices * gg.nang * gg.nrad;
        const int64_t nin = mode == 1 ? ns_ : nv, nout = mode == 0 ? ns_ : nv;
        const bool dev = flags & LC_DEVICE_PTRS;
        const float *din = in;
        float *dout = out;
        if (!dev) {
            float *buf = (float *)ctx->c.scratch(2, sizeof(float) * (size_t)(nin + nout));
            copy_in(ctx->c, buf, in, sizeof(float) * nin, false);
            din = buf;
            dout = buf + nin;
        }
        if (mode == 0) forward2d_dev(ctx->c, g->g, din, dout, (int)nslices);
        else if (mode == 1) backproj2d_dev(ctx->c, g->g, din, dout, (int)nslices);
        else sysmat2d_nufft_dev(ctx->c, g->g, din, dout, (int)nslices);
        if (!dev) copy_out(ctx->c, out, dout, sizeof(float) * nout, false);
        ctx->c.sync();
    });
}

int lc_forward2d(lc_ctx ctx, lc_geom2d g, const float *vol, int64_t nslices, float *sino,
                 uint32_t flags) {
    return project2d(ctx, g, vol, nslices, sino, flags, 0);
}

int lc_backproj2d(lc_ctx ctx, lc_geom2d g, const float *sino, int64_t nslices, float *vol,
                  uint32_t flags) {
    return project2d(ctx, g, sino, nslices, vol, flags, 1);
}

int lc_sysmat2d_nufft(lc_ctx ctx, lc_geom2d g, const float *x, int64_t nslices, float *y,
                      uint32_t flags) {
    return project2d(ctx, g, x, nslices, y, flags, 2);
}

} // extern "C"

template <typename T>
static int cg2d_t(lc_ctx ctx, lc_geom2d g, const T *rhs, T *x, int64_t nslices, int64_t max_iter,
                  double tol, double xtol, uint32_t flags, lc_report *rep) {
    return guard([&] {
        LC_REQUIRE(ctx && g && rhs && x, "lc_cg2d: NULL argument");
        LC_REQUIRE(nslices >= 1 && max_iter >= 0, "lc_cg2d: bad sizes");
        cg_one<T>(ctx, ProbSpec{&g->g, (int)nslices}, rhs, x, max_iter, tol, xtol, flags, rep);
    });
}

extern "C" {

int lc_cg2d(lc_ctx ctx, lc_geom2d g, const float *rhs, float *x, int64_t nslices,
            int64_t max_iter, double tol, double xtol, uint32_t flags, lc_report *rep) {
    return cg2d_t(ctx, g, rhs, x, nslices, max_iter, tol, xtol, flags, rep);
}

int lc_cg2d_f64(lc_ctx ctx, lc_geom2d g, const double *rhs, double *x, int64_t nslices,
                int64_t max_iter, double tol, double xtol, uint32_t flags, lc_report *rep) {
    return cg2d_t(ctx, g, rhs, x, nslices, max_iter, tol, xtol, flags, rep);
}

int lc_cg2d_noalloc(lc_ctx ctx, lc_geom2d g, const float *rhs, float *x, int64_t nslices,
                    int64_t max_iter, double tol, double xtol, uint32_t flags, lc_report *rep) {
    return guard([&] {
        LC_REQUIRE(ctx && g && rhs && x, "lc_cg2d_noalloc: NULL argument");
        LC_REQUIRE(nslices >= 1 && nslices <= (1 << 20), "lc_cg2d_noalloc: bad nslices");
        LC_REQUIRE(g->ctx == ctx, "lc_cg2d_noalloc: geometry of another context");
        const CgState s = cg2d_noalloc(ctx->c, g->g, rhs, x, (int)nslices, max_iter, tol, xtol,
                                       flags);
        if (rep) fill_report(s, rep);
    });
}

} // extern "C"

template <typename T>
static int sb2d_t(lc_ctx ctx, lc_geom2d g, const T *rhs, T *x, int64_t nslices, double lambda,
                  double mu, int64_t outer, int64_t inner, double tol, double xtol, uint32_t flags,
                  lc_report *rep) {
    return guard([&] {
        LC_REQUIRE(ctx && g && rhs && x, "lc_split_bregman2d: NULL argument");
        LC_REQUIRE(nslices >= 1 && outer >= 0 && inner >= 0, "lc_split_bregman2d: bad sizes");
        sb_one<T>(ctx, ProbSpec{&g->g, (int)nslices}, rhs, x, lambda, mu, outer, inner, tol, xtol,
                  flags, rep);
    });
}

extern "C" {

int lc_split_bregman2d(lc_ctx ctx, lc_geom2d g, const float *rhs, float *x, int64_t nslices,
                       double lambda, double mu, int64_t outer, int64_t inner, double tol,
                       double xtol, uint32_t flags, lc_report *rep) {
    return sb2d_t(ctx, g, rhs, x, nslices, lambda, mu, outer, inner, tol, xtol, flags, rep);
}

int lc_split_bregman2d_f64(lc_ctx ctx, lc_geom2d g, const double *rhs, double *x,
                           int64_t nslices, double lambda, double mu, int64_t outer,
                           int64_t inner, double tol, double xtol, uint32_t flags,
                           lc_report *rep) {
    return sb2d_t(ctx, g, rhs, x, nslices, lambda, mu, outer, inner, tol, xtol, flags, rep);
}

// ---- 3-D ----
int lc_geom3d_create(lc_ctx ctx, const double *theta, const double *gamma, const double *beta,
                     int64_t nproj, int64_t nrows, int64_t ncols, const int64_t *recon_dims,
                     uint32_t flags, lc_geom3d *out) {
    return guard([&] {
        LC_REQUIRE(ctx && theta && gamma && beta && recon_dims && out, "lc_geom3d_create: NULL argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        auto *g = new lc_geom3d_s();
        g->ctx = ctx;
        try {
            geom3d_init(ctx->c, g->g, theta, gamma, beta, nproj, nrows, ncols, recon_dims);
            if (!(flags & LC_GEOM_LAZY_PSF)) geom3d_build_psf(ctx->c, g->g);
            ctx->c.sync();
        } catch (...) {
            delete g;
            throw;
        }
        ctx_retain(ctx);
        *out = g;
    });
}

int lc_geom3d_destroy(lc_geom3d g) {
    return guard([&] { destroy_on_ctx(g); });
}

int lc_geom3d_info(lc_geom3d g, int64_t *info) {
    return guard([&] {
        LC_REQUIRE(g && info, "lc_geom3d_info: NULL argument");
        const Geom3D &q = g->g;
        const int64_t v[9] = {q.nproj, q.nrows, q.ncols, q.n1, q.n2, q.n3, q.M1, q.Mr, q.Mr};
        for (int i = 0; i < 9; ++i) info[i] = v[i];
    });
}

int lc_psf3d_kernel_hat(lc_ctx ctx, lc_geom3d g, float *out) {
    return guard([&] {
        LC_REQUIRE(ctx && g && out, "lc_psf3d_kernel_hat: NULL argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        geom3d_kernel_hat<float>(ctx->c, g->g, out);
    });
}

int lc_psf3d_kernel_hat_f64(lc_ctx ctx, lc_geom3d g, double *out) {
    return guard([&] {
        LC_REQUIRE(ctx && g && out, "lc_psf3d_kernel_hat_f64: NULL argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        geom3d_kernel_hat<double>(ctx->c, g->g, out);
    });
}

} // extern "C"

template <typename T>
static int sysmat3d_psf_t(lc_ctx ctx, lc_geom3d g, const T *x, T *y, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx && g && x && y, "lc_sysmat3d_psf: NULL argument");
        ProbSpec ps{nullptr, (int)g->g.n1, &g->g};
        sysmat_psf<T>(ctx, ps, x, y, flags);
    });
}

extern "C" {

int lc_sysmat3d_psf(lc_ctx ctx, lc_geom3d g, const float *x, float *y, uint32_t flags) {
    return sysmat3d_psf_t(ctx, g, x, y, flags);
}

int lc_sysmat3d_psf_f64(lc_ctx ctx, lc_geom3d g, const double *x, double *y, uint32_t flags) {
    return sysmat3d_psf_t(ctx, g, x, y, flags);
}

// forward (mode 0) / backproj (1) / NUFFT sysmat (2) on one volume
static int project3d(lc_ctx ctx, lc_geom3d g, const float *in, float *out, const double *shifts,
                     uint32_t flags, int mode) {
    return guard([&] {
        LC_REQUIRE(ctx && g && in && out, "3-D projector: NULL argument");
        LC_REQUIRE(g->ctx == ctx, "3-D projector: geometry of another context");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        const Geom3D &q = g->g;
        const int64_t nv = q.n1 * q.n2 * q.n3, np = q.nproj * q.nrows * q.ncols;
        const int64_t nin = mode == 1 ? np : nv, nout = mode == 0 ? np : nv;
        const bool dev = flags & LC_DEVICE_PTRS;
        const float *din = in;
        float *dout = out;
        if (!dev) {
            float *buf = (float *)ctx->c.scratch(2, sizeof(float) * (size_t)(nin + nout));
            copy_in(ctx->c, buf, in, sizeof(float) * nin, false);
            din = buf;
            dout = buf + nin;
        }
        DBuf dsh;
        if (mode == 1 && shifts) {
            dsh.alloc(sizeof(double) * 2 * q.nproj);
            copy_in(ctx->c, dsh.as(), shifts, sizeof(double) * 2 * q.nproj, dev);
        }
        if (mode == 0) forward3d_dev(ctx->c, g->g, din, dout);
        else if (mode == 1) backproj3d_dev(ctx->c, g->g, din, dout, shifts ? dsh.as<double>() : nullptr);
        else sysmat3d_nufft_dev(ctx->c, g->g, din, dout);
        if (!dev) copy_out(ctx->c, out, dout, sizeof(float) * nout, false);
        ctx->c.sync();
    });
}

int lc_forward3d(lc_ctx ctx, lc_geom3d g, const float *vol, float *proj, uint32_t flags) {
    return project3d(ctx, g, vol, proj, nullptr, flags, 0);
}

int lc_backproj3d(lc_ctx ctx, lc_geom3d g, const float *proj, const double *shifts, float *vol,
                  uint32_t flags) {
    return project3d(ctx, g, proj, vol, shifts, flags, 1);
}

int lc_sysmat3d_nufft(lc_ctx ctx, lc_geom3d g, const float *x, float *y, uint32_t flags) {
    return project3d(ctx, g, x, y, nullptr, flags, 2);
}

} // extern "C"

template <typename T>
static int cg3d_t(lc_ctx ctx, lc_geom3d g, const T *rhs, T *x, int64_t max_iter, double tol,
                  double xtol, uint32_t flags, lc_report *rep) {
    return guard([&] {
        LC_REQUIRE(ctx && g && rhs && x, "lc_cg3d: NULL argument");
        LC_REQUIRE(max_iter >= 0, "lc_cg3d: bad max_iter");
        cg_one<T>(ctx, ProbSpec{nullptr, (int)g->g.n1, &g->g}, rhs, x, max_iter, tol, xtol, flags,
                  rep);
    });
}

extern "C" {

int lc_cg3d(lc_ctx ctx, lc_geom3d g, const float *rhs, float *x, int64_t max_iter, double tol,
            double xtol, uint32_t flags, lc_report *rep) {
    return cg3d_t(ctx, g, rhs, x, max_iter, tol, xtol, flags, rep);
}

int lc_cg3d_f64(lc_ctx ctx, lc_geom3d g, const double *rhs, double *x, int64_t max_iter,
                double tol, double xtol, uint32_t flags, lc_report *rep) {
    return cg3d_t(ctx, g, rhs, x, max_iter, tol, xtol, flags, rep);
}

} // extern "C"

template <typename T>
static int sb3d_t(lc_ctx ctx, lc_geom3d g, const T *rhs, T *x, double lambda, double mu,
                  int64_t outer, int64_t inner, double tol, double xtol, uint32_t flags,
                  lc_report *rep) {
    return guard([&] {
        LC_REQUIRE(ctx && g && rhs && x, "lc_split_bregman3d: NULL argument");
        LC_REQUIRE(outer >= 0 && inner >= 0, "lc_split_bregman3d: bad sizes");
        sb_one<T>(ctx, ProbSpec{nullptr, (int)g->g.n1, &g->g}, rhs, x, lambda, mu, outer, inner,
                  tol, xtol, flags, rep);
    });
}

extern "C" {

int lc_split_bregman3d(lc_ctx ctx, lc_geom3d g, const float *rhs, float *x, double lambda,
                       double mu, int64_t outer, int64_t inner, double tol, double xtol,
                       uint32_t flags, lc_report *rep) {
    return sb3d_t(ctx, g, rhs, x, lambda, mu, outer, inner, tol, xtol, flags, rep);
}

int lc_split_bregman3d_f64(lc_ctx ctx, lc_geom3d g, const double *rhs, double *x, double lambda,
                           double mu, int64_t outer, int64_t inner, double tol, double xtol,
                           uint32_t flags, lc_report *rep) {
    return sb3d_t(ctx, g, rhs, x, lambda, mu, outer, inner, tol, xtol, flags, rep);
}

} // extern "C"

template <typename T>
static int fd_call(lc_ctx ctx, int op, const T *in, T *out, int64_t n1, int64_t n2, int64_t n3,
                   uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx && in && out, "finite differences: NULL argument");
        LC_REQUIRE(n1 >= 1 && n2 >= 1 && n3 >= 1, "finite differences: bad dims");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        const int64_t n = n1 * n2 * n3;
        const int64_t nin = (op == 1) ? 3 * n : n, nout = (op == 0) ? 3 * n : n;
        const bool dev = flags & LC_DEVICE_PTRS;
        const T *din = in;
        T *dout = out;
        if (!dev) {
            T *buf = (T *)ctx->c.scratch(2, sizeof(T) * (size_t)(nin + nout));
            copy_in(ctx->c, buf, in, sizeof(T) * nin, false);
            din = buf;
            dout = buf + nin;
        }
        fd_launch(ctx->c, op, din, dout, (int)n1, (int)n2, (int)n3,
                  (flags & LC_TV_ZERO_PAD) ? 1 : 0);
        if (!dev) copy_out(ctx->c, out, dout, sizeof(T) * nout, false);
        ctx->c.sync();
    });
}

extern "C" {

int lc_grad(lc_ctx ctx, const float *u, int64_t n1, int64_t n2, int64_t n3, float *g,
            uint32_t flags) {
    return fd_call(ctx, 0, u, g, n1, n2, n3, flags);
}

int lc_divergence(lc_ctx ctx, const float *g, int64_t n1, int64_t n2, int64_t n3, float *out,
                  uint32_t flags) {
    return fd_call(ctx, 1, g, out, n1, n2, n3, flags);
}

int lc_laplacian(lc_ctx ctx, const float *u, int64_t n1, int64_t n2, int64_t n3, float *out,
                 uint32_t flags) {
    return fd_call(ctx, 2, u, out, n1, n2, n3, flags);
}

int lc_grad_f64(lc_ctx ctx, const double *u, int64_t n1, int64_t n2, int64_t n3, double *g,
                uint32_t flags) {
    return fd_call(ctx, 0, u, g, n1, n2, n3, flags);
}

int lc_divergence_f64(lc_ctx ctx, const double *g, int64_t n1, int64_t n2, int64_t n3,
                      double *out, uint32_t flags) {
    return fd_call(ctx, 1, g, out, n1, n2, n3, flags);
}

int lc_laplacian_f64(lc_ctx ctx, const double *u, int64_t n1, int64_t n2, int64_t n3, double *out,
                     uint32_t flags) {
    return fd_call(ctx, 2, u, out, n1, n2, n3, flags);
}

int lc_batch2d_create(lc_ctx ctx, int32_t nprob, const lc_geom2d *geoms, const int32_t *ns,
                      lc_batch2d *out) {
    return lc_batch_create_ex(ctx, nprob, geoms, nullptr, ns, 0, out);
}

int lc_batch_create(lc_ctx ctx, int32_t nprob, const lc_geom2d *g2, const lc_geom3d *g3,
                    const int32_t *ns, lc_batch2d *out) {
    return lc_batch_create_ex(ctx, nprob, g2, g3, ns, 0, out);
}

int lc_batch_create_ex(lc_ctx ctx, int32_t nprob, const lc_geom2d *g2, const lc_geom3d *g3,
                       const int32_t *ns, uint32_t flags, lc_batch2d *out) {
    return guard([&] {
        LC_REQUIRE(ctx && out && nprob >= 1 && (g2 || g3), "lc_batch_create: bad argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        std::vector<ProbSpec> ps;
        for (int i = 0; i < nprob; ++i) {
            const lc_geom2d a = g2 ? g2[i] : nullptr;
            const lc_geom3d v = g3 ? g3[i] : nullptr;
            LC_REQUIRE((a != nullptr) != (v != nullptr),
                       "lc_batch_create: each problem needs exactly one geometry");
            if (a) {
                LC_REQUIRE(a->ctx == ctx, "lc_batch_create: geometry of another context");
                LC_REQUIRE(ns && ns[i] >= 1, "lc_batch_create: ns[i] must be >= 1 for a 2-D problem");
                ps.push_back(ProbSpec{&a->g, ns[i]});
            } else {
                LC_REQUIRE(v->ctx == ctx, "lc_batch_create: geometry of another context");
                ps.push_back(ProbSpec{nullptr, (int)v->g.n1, &v->g});
            }
        }
        auto *b = new lc_batch2d_s();
        b->ctx = ctx;
        try {
            b->b = std::make_unique<Batch2D>(ctx->c, ps, (flags & LC_FP64) != 0);
            // zero x and y
            const size_t bytes = b->b->esize() * b->b->total();
            LC_HIP(hipMemsetAsync(b->b->Xv(), 0, bytes, ctx->c.stream()));
            LC_HIP(hipMemsetAsync(b->b->Yv(), 0, bytes, ctx->c.stream()));
            ctx->c.sync();
        } catch (...) {
            delete b;
            throw;
        }
        b->reps.resize(nprob);
        ctx_retain(ctx);
        *out = b;
    });
}

int lc_batch2d_destroy(lc_batch2d b) {
    return guard([&] { destroy_on_ctx(b); });
}

} // extern "C"

template <typename T>
static int batch_set_t(lc_batch2d b, int32_t prob, int32_t which, const T *src, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(b && src && prob >= 0 && prob < b->b->nprob(), "lc_batch2d_set: bad argument");
        batch_check_type<T>(*b->b, "lc_batch2d_set");
        std::lock_guard<std::mutex> lk(b->ctx->c.mu);
        b->ctx->c.set_current();
        void *dst = b->b->elem(batch_buf(*b->b, which), b->b->offset(prob));
        copy_in(b->ctx->c, dst, src, sizeof(T) * b->b->size(prob), flags & LC_DEVICE_PTRS);
        b->ctx->c.sync();
    });
}

extern "C" {

} // extern "C"

template <typename T>
static int batch_get_t(lc_batch2d b, int32_t prob, int32_t which, T *dst, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(b && dst && prob >= 0 && prob < b->b->nprob(), "lc_batch2d_get: bad argument");
        batch_check_type<T>(*b->b, "lc_batch2d_get");
        std::lock_guard<std::mutex> lk(b->ctx->c.mu);
        b->ctx->c.set_current();
        const void *src = b->b->elem(batch_buf(*b->b, which), b->b->offset(prob));
        copy_out(b->ctx->c, dst, src, sizeof(T) * b->b->size(prob), flags & LC_DEVICE_PTRS);
        b->ctx->c.sync();
    });
}

extern "C" {

} // extern "C"

template <typename T>
static int batch_ptr_t(lc_batch2d b, int32_t prob, int32_t which, T **out) {
    return guard([&] {
        LC_REQUIRE(b && out && prob >= 0 && prob < b->b->nprob(), "lc_batch2d_device_ptr: bad argument");
        batch_check_type<T>(*b->b, "lc_batch2d_device_ptr");
        *out = static_cast<T *>(b->b->elem(batch_buf(*b->b, which), b->b->offset(prob)));
    });
}

extern "C" {

int lc_batch2d_set(lc_batch2d b, int32_t prob, int32_t which, const float *src, uint32_t flags) {
    return batch_set_t(b, prob, which, src, flags);
}
int lc_batch2d_get(lc_batch2d b, int32_t prob, int32_t which, float *dst, uint32_t flags) {
    return batch_get_t(b, prob, which, dst, flags);
}
int lc_batch2d_device_ptr(lc_batch2d b, int32_t prob, int32_t which, float **out) {
    return batch_ptr_t(b, prob, which, out);
}
int lc_batch2d_set_f64(lc_batch2d b, int32_t prob, int32_t which, const double *src,
                       uint32_t flags) {
    return batch_set_t(b, prob, which, src, flags);
}
int lc_batch2d_get_f64(lc_batch2d b, int32_t prob, int32_t which, double *dst, uint32_t flags) {
    return batch_get_t(b, prob, which, dst, flags);
}
int lc_batch2d_device_ptr_f64(lc_batch2d b, int32_t prob, int32_t which, double **out) {
    return batch_ptr_t(b, prob, which, out);
}

int lc_batch2d_cg(lc_batch2d b, const int32_t *max_iter, double tol, double xtol,
                  uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(b && max_iter, "lc_batch2d_cg: NULL argument");
        std::lock_guard<std::mutex> lk(b->ctx->c.mu);
        b->ctx->c.set_current();
        std::vector<int> mi(max_iter, max_iter + b->b->nprob());
        const bool graph = flags & LC_GRAPH, hsync = (flags & LC_HOST_SYNC) && !graph;
        b->b->cg(mi, tol, xtol, graph, hsync, (flags & LC_X0_ZERO) != 0);
        auto st = b->b->cg_state();
        for (int i = 0; i < b->b->nprob(); ++i) fill_report(st[i], &b->reps[i]);
    });
}

int lc_batch2d_split_bregman(lc_batch2d b, const float *lambda, const float *mu, int64_t outer,
                             int64_t inner, double tol, double xtol, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(b && lambda && mu, "lc_batch2d_split_bregman: NULL argument");
        std::lock_guard<std::mutex> lk(b->ctx->c.mu);
        b->ctx->c.set_current();
        const int np = b->b->nprob();
        std::vector<float> la(lambda, lambda + np), m(mu, mu + np);
        const bool graph = flags & LC_GRAPH, hsync = (flags & LC_HOST_SYNC) && !graph;
        b->b->split_bregman(la, m, (int)outer, (int)inner, tol, xtol,
                            (flags & LC_TV_ZERO_PAD) ? 1 : 0, (flags & LC_TV_ANISO) != 0, graph,
                            hsync, (flags & LC_X0_ZERO) != 0);
        auto st = b->b->cg_state();
        auto sb = b->b->sb_state();
        for (int i = 0; i < np; ++i) {
            fill_report(st[i], &b->reps[i]);
            b->reps[i].outer = sb[i].outer;
            b->reps[i].dx = sb[i].norm_diff;
        }
    });
}

int lc_batch2d_reports(lc_batch2d b, lc_report *reps) {
    return guard([&] {
        LC_REQUIRE(b && reps, "lc_batch2d_reports: NULL argument");
        std::memcpy(reps, b->reps.data(), sizeof(lc_report) * b->reps.size());
    });
}

int lc_autotune2d(lc_ctx ctx, lc_geom2d g, int32_t nslices, int32_t *variant, double *ms) {
    return guard([&] {
        LC_REQUIRE(ctx && g && nslices >= 1, "lc_autotune2d: bad argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        double t = 0.0;
        const ToepVariant v = toep_autotune(ctx->c, g->g, nslices, &t);
        if (variant) {
            variant[0] = v.row;
            variant[1] = v.col;
            variant[2] = v.B;
        }
        if (ms) *ms = t;
    });
}

int lc_set_tuning2d(lc_ctx ctx, int64_t image_n, const int32_t *variant) {
    return guard([&] {
        LC_REQUIRE(ctx && variant && image_n >= 2, "lc_set_tuning2d: bad argument");
        LC_REQUIRE(variant[0] >= 0 && variant[0] < toep_num_row_shapes() && variant[1] >= 0 &&
                       variant[1] < toep_num_col_shapes() && variant[2] >= 0 && variant[2] <= 16,
                   "lc_set_tuning2d: variant out of range");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ToepVariant v;
        v.row = variant[0];
        v.col = variant[1];
        v.B = variant[2];
        ctx->c.toep_tuning[(int)image_n] = v;
    });
}

int lc_get_tuning2d(lc_ctx ctx, int64_t image_n, int32_t *variant) {
    return guard([&] {
        LC_REQUIRE(ctx && variant, "lc_get_tuning2d: bad argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        auto it = ctx->c.toep_tuning.find((int)image_n);
        ToepVariant v = it != ctx->c.toep_tuning.end()
                            ? it->second
                            : toep_default_variant((int)toep_conv_size(image_n));
        variant[0] = v.row;
        variant[1] = v.col;
        variant[2] = v.B;
    });
}

int lc_autotune2d_nufft(lc_ctx ctx, lc_geom2d g, int32_t nslices, int32_t *shape, double *ms) {
    return guard([&] {
        LC_REQUIRE(ctx && g && nslices >= 1 && nslices <= 65535, "lc_autotune2d_nufft: bad argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        double t = 0.0;
        const int i = nufft2d_autotune(ctx->c, g->g, nslices, &t);
        if (shape) *shape = i;
        if (ms) *ms = t;
    });
}

int lc_set_tuning2d_nufft(lc_ctx ctx, lc_geom2d g, int32_t shape) {
    return guard([&] {
        LC_REQUIRE(ctx && g, "lc_set_tuning2d_nufft: bad argument");
        LC_REQUIRE(shape >= 0 && shape < nufft2d_num_shapes(), "lc_set_tuning2d_nufft: shape out of range");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.nu2_tuning[nufft2d_fine_n(g->g.N)] = shape;
    });
}

int lc_get_tuning2d_nufft(lc_ctx ctx, lc_geom2d g, int32_t *shape) {
    return guard([&] {
        LC_REQUIRE(ctx && g && shape, "lc_get_tuning2d_nufft: bad argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        *shape = ctx->c.nu2_shape(nufft2d_fine_n(g->g.N));
    });
}

int lc_autotune3d(lc_ctx ctx, lc_geom3d g, int32_t *shape, double *ms) {
    return guard([&] {
        LC_REQUIRE(ctx && g, "lc_autotune3d: bad argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx->c.set_current();
        double t = 0.0;
        const int i = nufft3d_autotune(ctx->c, g->g, &t);
        if (shape) *shape = i;
        if (ms) *ms = t;
    });
}

int lc_set_tuning3d(lc_ctx ctx, lc_geom3d g, int32_t shape) {
    return guard([&] {
        LC_REQUIRE(ctx && g, "lc_set_tuning3d: bad argument");
        LC_REQUIRE(shape >= 0 && shape < nufft3d_num_shapes(), "lc_set_tuning3d: shape out of range");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        const int64_t n[3] = {g->g.n1, g->g.n2, g->g.n3};
        ctx->c.nu3_tuning[{nufft3d_fine_n(n[0]), nufft3d_fine_n(n[1]), nufft3d_fine_n(n[2])}] = shape;
    });
}

int lc_get_tuning3d(lc_ctx ctx, lc_geom3d g, int32_t *shape) {
    return guard([&] {
        LC_REQUIRE(ctx && g && shape, "lc_get_tuning3d: bad argument");
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        const int nf[3] = {nufft3d_fine_n(g->g.n1), nufft3d_fine_n(g->g.n2), nufft3d_fine_n(g->g.n3)};
        *shape = ctx->c.nu3_shape(nf);
    });
}

int lc_batch2d_profile(lc_batch2d b, int32_t reps, int32_t max_rows, double *rows,
                       int32_t *nrows) {
    return guard([&] {
        LC_REQUIRE(b && rows && nrows && reps >= 1, "lc_batch2d_profile: bad argument");
        std::lock_guard<std::mutex> lk(b->ctx->c.mu);
        b->ctx->c.set_current();
        const auto t = b->b->profile_passes(reps);
        LC_REQUIRE((int)t.size() <= max_rows, "lc_batch2d_profile: max_rows too small");
        for (size_t i = 0; i < t.size(); ++i) {
            double *r = rows + 5 * i;
            r[0] = t[i].N;
            r[1] = t[i].pass;
            r[2] = t[i].nslices;
            r[3] = t[i].ms;
            r[4] = t[i].bytes;
        }
        *nrows = (int32_t)t.size();
    });
}

int lc_batch2d_sysmat(lc_batch2d b, int32_t src, int32_t dst) {
    return guard([&] {
        LC_REQUIRE(b, "lc_batch2d_sysmat: NULL batch");
        LC_REQUIRE(src != dst, "lc_batch2d_sysmat: src and dst must differ");
        std::lock_guard<std::mutex> lk(b->ctx->c.mu);
        b->ctx->c.set_current();
        b->b->toeplitz(batch_buf(*b->b, src), batch_buf(*b->b, dst));
    });
}

// ---- server ----
struct lc_server_s {
    lc_ctx_s *ctx;
    std::unique_ptr<Server> s;
};

int lc_server_create(lc_ctx ctx, uint64_t hbm_budget, int32_t lanes, int32_t max_batch,
                     lc_server *out) {
    return guard([&] {
        LC_REQUIRE(ctx && out && hbm_budget > 0, "lc_server_create: bad argument");
        auto *sv = new lc_server_s();
        sv->ctx = ctx;
        try {
            sv->s = std::make_unique<Server>(ctx->c, (size_t)hbm_budget, lanes, max_batch);
        } catch (...) {
            delete sv;
            throw;
        }
        ctx_retain(ctx);
        *out = sv;
    });
}

int lc_server_submit(lc_server s, const lc_request *r, uint64_t *ticket) {
    return guard([&] {
        LC_REQUIRE(s && r && ticket, "lc_server_submit: NULL argument");
        LC_REQUIRE((r->g2 != nullptr) != (r->g3 != nullptr), "lc_server_submit: need exactly one geometry");
        LC_REQUIRE(r->sino || r->rhs, "lc_server_submit: need a sinogram or a right-hand side");
        LC_REQUIRE(r->solver == 0 || r->solver == 1, "lc_server_submit: solver must be 0 (CG) or 1 (SB)");
        auto q = std::make_unique<ServerRequest>();
        int64_t nvol, nsino;
        if (r->g3) {
            LC_REQUIRE(r->g3->ctx == s->ctx, "lc_server_submit: geometry of another context");
            const Geom3D &g = r->g3->g;
            q->g3 = &r->g3->g;
            q->ns = (int)g.n1;
            nvol = g.n1 * g.n2 * g.n3;
            nsino = g.nproj * g.nrows * g.ncols;
        } else {
            LC_REQUIRE(r->g2->ctx == s->ctx, "lc_server_submit: geometry of another context");
            LC_REQUIRE(r->nslices >= 1, "lc_server_submit: nslices must be >= 1");
            const Geom2D &g = r->g2->g;
            q->g2 = &r->g2->g;
            q->ns = r->nslices;
            nvol = (int64_t)r->nslices * g.N * g.N;
            nsino = (int64_t)r->nslices * g.nang * g.nrad;
        }
        q->solver = r->solver;
        q->max_iter = r->max_iter;
        q->outer = r->outer;
        q->inner = r->inner;
        LC_REQUIRE(q->max_iter >= 0 && q->outer >= 0 && q->inner >= 0, "lc_server_submit: negative iterations");
        q->tol = r->tol;
        q->xtol = r->xtol;
        q->lambda = (float)r->lambda;
        q->mu = (float)r->mu;
        q->flags = r->flags;
        const bool borrow = (r->flags & LC_REQ_BORROW) != 0;
        q->has_sino = r->sino != nullptr;
        q->in_n = (size_t)(r->sino ? nsino : nvol);
        const float *in = r->sino ? r->sino : r->rhs;
        if (borrow) {
            q->in = in;
        } else {
            (r->sino ? q->sino : q->rhs).assign(in, in + q->in_n);
            q->in = (r->sino ? q->sino : q->rhs).data();
        }
        if (r->x0) {
            q->x0_n = (size_t)nvol;
            if (borrow) {
                q->x0p = r->x0;
            } else {
                q->x0.assign(r->x0, r->x0 + nvol);
                q->x0p = q->x0.data();
            }
        }
        q->x_out = r->x_out;
        *ticket = s->s->submit(std::move(q));
    });
}

int lc_server_poll(lc_server s, uint64_t ticket, int32_t *done) {
    return guard([&] {
        LC_REQUIRE(s && done, "lc_server_poll: NULL argument");
        *done = s->s->poll(ticket) ? 1 : 0;
    });
}

int lc_server_wait(lc_server s, uint64_t ticket, double timeout_s, int32_t *done) {
    return guard([&] {
        LC_REQUIRE(s, "lc_server_wait: NULL server");
        const bool d = s->s->wait(ticket, timeout_s);
        if (done) *done = d ? 1 : 0;
    });
}

int lc_server_result(lc_server s, uint64_t ticket, float *x, lc_report *rep, double *times) {
    return guard([&] {
        LC_REQUIRE(s, "lc_server_result: NULL server");
        auto r = s->s->take(ticket);
        if (r->state == 3) throw Error(LC_EINTERNAL, "server: request failed: " + r->err);
        if (!r->x_out) {
            LC_REQUIRE(x, "lc_server_result: NULL x for a request without x_out");
            std::memcpy(x, r->x.data(), sizeof(float) * r->x.size());
        } else if (x && x != r->x_out) {
            std::memcpy(x, r->x_out, sizeof(float) * (size_t)(r->g3 ? r->g3->n1 * r->g3->n2 * r->g3->n3
                                                                    : (int64_t)r->ns * r->g2->N * r->g2->N));
        }
        if (rep) *rep = r->rep;
        if (times) {
            times[0] = r->t_submit;
            times[1] = r->t_start;
            times[2] = r->t_done;
        }
    });
}

int lc_server_stats(lc_server s, double *stats) {
    return guard([&] {
        LC_REQUIRE(s && stats, "lc_server_stats: NULL argument");
        s->s->stats(stats);
    });
}

int lc_server_timeline(lc_server s, int32_t max_rows, double *rows, int32_t *nrows) {
    return guard([&] {
        LC_REQUIRE(s && nrows && max_rows >= 0 && (rows || max_rows == 0),
                   "lc_server_timeline: bad argument");
        const std::vector<BatchRecord> v = s->s->timeline((size_t)max_rows);
        for (size_t i = 0; i < v.size(); ++i) {
            double *o = rows + 6 * i;
            o[0] = v[i].lane;
            o[1] = v[i].nreq;
            o[2] = v[i].h0;
            o[3] = v[i].h1;
            o[4] = v[i].d0;
            o[5] = v[i].d1;
        }
        *nrows = (int32_t)v.size();
    });
}

int lc_reference_memory_model(const int64_t *dims, int32_t elem_bytes, int32_t regularizer,
                              double *out) {
    return guard([&] {
        LC_REQUIRE(dims && out, "lc_reference_memory_model: NULL argument");
        LC_REQUIRE(elem_bytes == 4 || elem_bytes == 8, "lc_reference_memory_model: elem_bytes 4 or 8");
        LC_REQUIRE(regularizer == LC_REG_SPLIT_BREGMAN || regularizer == LC_REG_UNCONSTRAINED,
                   "lc_reference_memory_model: unknown regularizer");
        LC_REQUIRE(dims[0] >= 0 && dims[1] >= 0 && dims[2] >= 0, "lc_reference_memory_model: negative dims");
        const double N = (double)dims[0] * (double)dims[1] * (double)dims[2];
        const bool sb = regularizer == LC_REG_SPLIT_BREGMAN;
        const double e = (double)elem_bytes;
        out[0] = (sb ? 22.0 : 11.0) * N * e;
        out[1] = (sb ? 24.0 : 11.0) * N * e;
        out[2] = 2.0 * N * e;
        out[3] = (sb ? 27.0 : 11.0) * N * e;
        out[4] = N;
    });
}

int lc_volume_fits(const lc_recon_params *p, int32_t elem_bytes, uint64_t free_bytes, int32_t *fits) {
    return guard([&] {
        LC_REQUIRE(p && fits, "lc_volume_fits: NULL argument");
        LC_REQUIRE(elem_bytes == 4 || elem_bytes == 8, "lc_volume_fits: elem_bytes 4 or 8");
        const uint64_t N = (uint64_t)p->recon_dims[0] * (uint64_t)p->recon_dims[1] *
                           (uint64_t)p->recon_dims[2];
        const uint64_t factor = p->regularizer == LC_REG_SPLIT_BREGMAN ? 27 : 11;
        // size_t arithmetic and the truncating cast of the reference's rule
        *fits = factor * N * (uint64_t)elem_bytes < (uint64_t)(0.8 * (double)free_bytes) ? 1 : 0;
    });
}

int lc_volume_fits_one_gpu(lc_ctx ctx, const lc_recon_params *p, int32_t elem_bytes, int32_t *fits) {
    return guard([&] {
        LC_REQUIRE(ctx && p && fits, "lc_volume_fits_one_gpu: NULL argument");
        ctx->c.set_current();
        size_t free_b = 0, total_b = 0;
        LC_HIP(hipMemGetInfo(&free_b, &total_b));
        const int st = lc_volume_fits(p, elem_bytes, (uint64_t)free_b, fits);
        if (st != LC_OK) throw Error(st, lc_last_error());
    });
}

int lc_request_bytes(const lc_request *r, uint64_t *bytes) {
    return guard([&] {
        LC_REQUIRE(r && bytes, "lc_request_bytes: NULL argument");
        LC_REQUIRE((r->g2 != nullptr) != (r->g3 != nullptr), "lc_request_bytes: need exactly one geometry");
        ServerRequest q;
        if (r->g3) {
            q.g3 = &r->g3->g;
            q.ns = (int)q.g3->n1;
        } else {
            LC_REQUIRE(r->nslices >= 1, "lc_request_bytes: nslices must be >= 1");
            q.g2 = &r->g2->g;
            q.ns = r->nslices;
        }
        q.solver = r->solver;
        q.has_sino = r->sino != nullptr;
        *bytes = (uint64_t)request_bytes(q);
    });
}

int lc_server_destroy(lc_server s) {
    return guard([&] {
        if (!s) return;
        lc_ctx_s *c = s->ctx;
        delete s;
        ctx_release(c);
    });
}

} // extern "C"

// ---- whole-solve drivers ----
extern "C" {

void lc_recon_params_default(lc_recon_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->regularizer = LC_REG_UNCONSTRAINED;
    p->max_iters = 100;
    p->inner_iters = 1;
    p->lambda = 0.1f;
    p->mu = 10.0f;
    p->tol = 1e-5f;
    p->xtol = 1e-5f;
    p->pad_factor = 1.4142f;
}

int lc_prepare_sinogram(lc_ctx ctx, int32_t ndatasets, const float *const *projs,
                        const int64_t *nang, int64_t nrows, int64_t ncols, float *sino,
                        float *max_out, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx, "lc_prepare_sinogram: NULL context");
        prepare_sinogram<float>(ctx->c, ndatasets, projs, nang, nrows, ncols, sino, max_out, flags);
    });
}

int lc_prepare_sinogram_f64(lc_ctx ctx, int32_t ndatasets, const double *const *projs,
                            const int64_t *nang, int64_t nrows, int64_t ncols, double *sino,
                            double *max_out, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx, "lc_prepare_sinogram_f64: NULL context");
        prepare_sinogram<double>(ctx->c, ndatasets, projs, nang, nrows, ncols, sino, max_out, flags);
    });
}

int lc_mbir2d(lc_ctx ctx, const float *sino, int64_t nslices, int64_t nang, int64_t ncols,
              const double *theta, const lc_recon_params *params, float *recon, lc_report *rep,
              uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx, "lc_mbir2d: NULL context");
        mbir2d<float>(ctx->c, sino, nslices, nang, ncols, theta, params, recon, rep, flags);
    });
}

int lc_mbir2d_f64(lc_ctx ctx, const double *sino, int64_t nslices, int64_t nang, int64_t ncols,
                  const double *theta, const lc_recon_params *params, double *recon,
                  lc_report *rep, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx, "lc_mbir2d_f64: NULL context");
        mbir2d<double>(ctx->c, sino, nslices, nang, ncols, theta, params, recon, rep, flags);
    });
}

int lc_mbir3d(lc_ctx ctx, int32_t ndatasets, const float *const *projs, const int64_t *nang,
              int64_t nrows, int64_t ncols, const double *theta, const double *gamma,
              const double *beta, const double *shifts, const lc_recon_params *params,
              float *recon, int64_t *padded_dims, lc_report *rep, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx, "lc_mbir3d: NULL context");
        mbir3d<float>(ctx->c, ndatasets, projs, nang, nrows, ncols, theta, gamma, beta, shifts, params,
                      recon, padded_dims, rep, flags);
    });
}

int lc_mbir3d_f64(lc_ctx ctx, int32_t ndatasets, const double *const *projs,
                  const int64_t *nang, int64_t nrows, int64_t ncols, const double *theta,
                  const double *gamma, const double *beta, const double *shifts,
                  const lc_recon_params *params, double *recon, int64_t *padded_dims,
                  lc_report *rep, uint32_t flags) {
    return guard([&] {
        LC_REQUIRE(ctx, "lc_mbir3d_f64: NULL context");
        mbir3d<double>(ctx->c, ndatasets, projs, nang, nrows, ncols, theta, gamma, beta, shifts, params,
                       recon, padded_dims, rep, flags);
    });
}

} // extern "C"

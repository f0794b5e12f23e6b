// Reconstruction request server: HBM-budget admission and continuous
// batching over several device lanes.
//
// A request is one reconstruction: a 2-D stack or a 3-D volume on a
// geometry, its sinogram (back-projected on the device) or right-hand side,
// and a CG or Split-Bregman solve.  Requests queue on the host.  Each lane is
// a context of its own (one stream, its own scratch) driven by one thread:
// whenever a lane is idle it takes every queued request that fits the free
// HBM budget and shares its solver settings, builds one batch, and runs the
// solve from captured graphs.  Requests arriving while other lanes are busy
// start on the next idle lane, so their work overlaps on the device with the
// batches already running.  A batch's results are copied to the callers'
// host arrays by a worker while its lane goes on to the next batch.  Each
// problem keeps its own CG / Bregman scalars,
// so its result does not depend on the batch it ran in.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>

#include "lc_engine.hpp"
#include "lc_server.hpp"

namespace lc {

namespace {
double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
} // namespace

// device bytes one request holds while it runs (batch vectors, Toeplitz
// intermediates, Split-Bregman fields, back-projection scratch)
size_t request_bytes(const ServerRequest &r) {
    int64_t N, ns, M, H, extra = 0, bp = 0;
    if (r.g3) {
        const Geom3D &g = *r.g3;
        N = g.N;
        ns = g.n1;
        M = g.M;
        H = g.H;
        extra = 8 * (H + 16) * M * ns;   // W: column blocks padded to B <= 16
        if (r.has_sino) {
            const Nufft3D &u = g.nu;
            int64_t nf[3];
            for (int a = 0; a < 3; ++a)
                nf[a] = u.ready ? u.nf[a] : nufft3d_fine_n(a == 0 ? g.n1 : a == 1 ? g.n2 : g.n3);
            // detector spectra (and the staged projections), the fine grid,
            // the line passes' intermediates and the spread's own buffers
            const int nfi[3] = {(int)nf[0], (int)nf[1], (int)nf[2]};
            const int64_t npts = g.nproj * g.nrows * g.ncols;
            bp = 8 * npts * 2 + 8 * (nf[0] * nf[1] * nf[2] + nf[0] * nf[1] * g.n3 + nf[0] * g.n2 * g.n3) +
                 (int64_t)nufft3d_spread_bytes(npts, nfi);
        }
    } else {
        const Geom2D &g = *r.g2;
        N = g.N;
        ns = r.ns;
        M = g.M;
        H = g.H;
        if (r.has_sino) {
            const int64_t nf = g.nu.ready ? g.nu.nf : nufft2d_fine_n(g.N);
            bp = 4 * ns * g.nang * g.nrad + 8 * ns * (g.nang * g.nrad + nf * nf + nf * N) +
                 (int64_t)nufft2d_spread_bytes(ns, g.nang * g.nrad, (int)nf);
        }
    }
    const int64_t n = ns * N * N;
    int64_t b = 5 * 4 * n + 8 * ns * (H + 16) * N + extra + bp;
    if (r.solver == 1) b += 4 * 8 * n;   // d, b (3 components each), x_old, y0
    return (size_t)b + (1 << 20);
}

Server::Server(Ctx &main, size_t budget, int nlanes, int max_batch)
    : main_(main), budget_(budget), max_batch_(max_batch) {
    LC_REQUIRE(nlanes >= 1 && nlanes <= 16, "server: 1 to 16 lanes");
    LC_REQUIRE(max_batch >= 1, "server: max_batch must be >= 1");
    main.set_current();
    LC_HIP(hipEventCreate(&epoch_));
    LC_HIP(hipEventRecord(epoch_, main.stream()));
    LC_HIP(hipEventSynchronize(epoch_));
    for (int i = 0; i < nlanes; ++i) {
        auto l = std::make_unique<Lane>();
        l->index = i;
        l->ctx = std::make_unique<Ctx>(main.device());
        l->ctx->toep_tuning = main.toep_tuning;
        l->ctx->nu3_tuning = main.nu3_tuning;
        l->ctx->nu2_tuning = main.nu2_tuning;
        LC_HIP(hipEventCreate(&l->ev0));
        LC_HIP(hipEventCreate(&l->ev1));
        lanes_.push_back(std::move(l));
    }
    for (auto &l : lanes_) l->th = std::thread([this, lp = l.get()] { lane_loop(*lp); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &l : lanes_) {
        if (l->th.joinable()) l->th.join();
        if (l->ev0) (void)hipEventDestroy(l->ev0);
        if (l->ev1) (void)hipEventDestroy(l->ev1);
    }
    if (epoch_) (void)hipEventDestroy(epoch_);
}

std::vector<BatchRecord> Server::timeline(size_t max_rows) {
    std::lock_guard<std::mutex> lk(m_);
    const size_t n = std::min(max_rows, records_.size());
    return std::vector<BatchRecord>(records_.end() - (ptrdiff_t)n, records_.end());
}

uint64_t Server::submit(std::unique_ptr<ServerRequest> r) {
    // geometry state the lanes only read: PSF and NUFFT tables, built here
    {
        std::lock_guard<std::mutex> lk(main_.mu);
        const bool build = r->g3 ? !r->g3->psf_ready || (r->has_sino && !r->g3->nu.ready)
                                 : !r->g2->psf_ready || (r->has_sino && !r->g2->nu.ready);
        if (build) {   // (a request on a ready geometry launches nothing here)
            // no lane captures while this thread allocates and synchronises
            std::lock_guard<std::recursive_mutex> cl(capture_mutex());
            main_.set_current();
            if (r->g3) {
                if (!r->g3->psf_ready) geom3d_build_psf(main_, *r->g3);
                if (r->has_sino && !r->g3->nu.ready) nufft3d_init(main_, *r->g3);
            } else {
                if (!r->g2->psf_ready) geom2d_build_psf(main_, *r->g2);
                if (r->has_sino && !r->g2->nu.ready) nufft2d_init(main_, *r->g2);
            }
            main_.sync();
        }
    }
    r->bytes = request_bytes(*r);
    LC_REQUIRE(r->bytes <= budget_, "server: request needs " + std::to_string(r->bytes) +
                                        " bytes, above the HBM budget of " + std::to_string(budget_));
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lk(m_);
    const uint64_t id = ++next_id_;
    r->id = id;
    queue_.push_back(id);
    reqs_[id] = std::move(r);
    ++queued_;
    cv_work_.notify_one();
    return id;
}

bool Server::wait(uint64_t id, double timeout_s) {
    std::unique_lock<std::mutex> lk(m_);
    auto it = reqs_.find(id);
    LC_REQUIRE(it != reqs_.end(), "server: unknown ticket");
    ServerRequest *r = it->second.get();
    auto done = [r] { return r->state >= 2; };
    if (timeout_s < 0) cv_done_.wait(lk, done);
    else cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    return r->state >= 2;
}

bool Server::poll(uint64_t id) {
    std::lock_guard<std::mutex> lk(m_);
    auto it = reqs_.find(id);
    LC_REQUIRE(it != reqs_.end(), "server: unknown ticket");
    return it->second->state >= 2;
}

std::unique_ptr<ServerRequest> Server::take(uint64_t id) {
    std::lock_guard<std::mutex> lk(m_);
    auto it = reqs_.find(id);
    LC_REQUIRE(it != reqs_.end(), "server: unknown ticket");
    LC_REQUIRE(it->second->state >= 2, "server: request not finished");
    auto r = std::move(it->second);
    reqs_.erase(it);
    return r;
}

void Server::stats(double *out) {
    std::lock_guard<std::mutex> lk(m_);
    out[0] = (double)inflight_;
    out[1] = (double)queued_;
    out[2] = (double)running_;
    out[3] = (double)done_;
    out[4] = (double)batches_;
    out[5] = (double)budget_;
    out[6] = t_build_;
    out[7] = t_input_;
    out[8] = t_solve_;
    out[9] = t_output_;
}

// requests that can share one batch solve
static bool same_group(const ServerRequest &a, const ServerRequest &b) {
    if (a.solver != b.solver || a.tol != b.tol || a.xtol != b.xtol) return false;
    if (a.solver == 1)
        return a.outer == b.outer && a.inner == b.inner &&
               (a.flags & (LC_TV_ZERO_PAD | LC_TV_ANISO)) == (b.flags & (LC_TV_ZERO_PAD | LC_TV_ANISO));
    return true;
}

// a head request that does not fit the free budget lets smaller requests
// start past it for at most this many batches; then the lanes wait until it fits
constexpr int kMaxBypass = 8;

void Server::lane_loop(Lane &lane) {
    tl_pool = &lane.pool;   // this lane's batches reuse device blocks instead of hipFree
    for (;;) {
        std::vector<ServerRequest *> batch;
        {
            std::unique_lock<std::mutex> lk(m_);
            for (;;) {
                if (stop_ && queue_.empty()) {
                    // (the copy-out takes m_ to mark its requests done)
                    lk.unlock();
                    if (lane.out_job.valid()) lane.out_job.get();
                    return;
                }
                if (!queue_.empty()) {
                    // the batch's anchor: the head when it fits, else the first
                    // queued request that fits (unless the head has been passed
                    // over kMaxBypass times); then every first-come request of
                    // the anchor's group that fits the remaining budget
                    size_t room = budget_ - inflight_;
                    ServerRequest *head = reqs_[queue_.front()].get();
                    ServerRequest *anchor = nullptr;
                    if (head->bytes <= room) {
                        anchor = head;
                    } else if (head->bypassed < kMaxBypass) {
                        for (uint64_t id : queue_) {
                            ServerRequest *r = reqs_[id].get();
                            if (r->bytes <= room) {
                                anchor = r;
                                break;
                            }
                        }
                    }
                    if (anchor) {
                        if (anchor != head) ++head->bypassed;
                        for (auto it = queue_.begin(); it != queue_.end() && (int)batch.size() < max_batch_;) {
                            ServerRequest *r = reqs_[*it].get();
                            if (same_group(*anchor, *r) && r->bytes <= room) {
                                room -= r->bytes;
                                batch.push_back(r);
                                it = queue_.erase(it);
                            } else {
                                ++it;
                            }
                        }
                        break;
                    }
                }
                cv_work_.wait(lk);
            }
            for (ServerRequest *r : batch) {
                inflight_ += r->bytes;
                r->state = 1;
                r->t_start = now_s();
            }
            queued_ -= batch.size();
            running_ += batch.size();
            ++batches_;
        }
        std::map<ServerRequest *, std::string> errs;
        std::vector<OutCopy> outs;
        bool isolated = false;
        try {
            outs = run_batch(lane, batch);
        } catch (const std::exception &e) {
            if (batch.size() == 1) {
                errs[batch[0]] = e.what();
            } else {
                // the batch failed: each request runs alone, so only a request
                // that fails on its own fails; its result is delivered at once
                // (the next run reuses the result staging)
                isolated = true;
                for (ServerRequest *r : batch) {
                    std::vector<ServerRequest *> one{r};
                    try {
                        const std::vector<OutCopy> o = run_batch(lane, one);
                        deliver(o, errs);
                    } catch (const std::exception &e1) {
                        errs[r] = e1.what();
                    }
                }
            }
        }
        {
            // the batch's device memory is released: other lanes may admit now
            std::lock_guard<std::mutex> lk(m_);
            for (ServerRequest *r : batch) inflight_ -= r->bytes;
        }
        cv_work_.notify_all();
        if (isolated || !errs.empty()) {
            finish(batch, errs);
            continue;
        }
        // the host copies of the results run beside the lane's next batch
        lane.out_job = std::async(std::launch::async, [this, batch, outs = std::move(outs)] {
            std::map<ServerRequest *, std::string> e;
            deliver(outs, e);
            finish(batch, e);
        });
    }
}

void Server::deliver(const std::vector<OutCopy> &outs, std::map<ServerRequest *, std::string> &errs) {
    for (const OutCopy &o : outs) {
        float *xo = o.r->x_out;
        if (!xo) {
            o.r->x.resize(o.n);
            xo = o.r->x.data();
        }
        std::memcpy(xo, o.src, sizeof(float) * o.n);
        size_t bad = 0;
        for (size_t i = 0; i < o.n; ++i) bad += !std::isfinite(xo[i]);
        if (bad)
            errs[o.r] = "the solution has " + std::to_string(bad) + " non-finite values of " +
                        std::to_string(o.n);
    }
}

// marks a batch's requests done (or failed, those with a message in errs)
// and wakes their waiters
void Server::finish(const std::vector<ServerRequest *> &batch, const std::map<ServerRequest *, std::string> &errs) {
    {
        std::lock_guard<std::mutex> lk(m_);
        const double t = now_s();
        for (ServerRequest *r : batch) {
            r->t_done = t;
            auto it = errs.find(r);
            r->state = it == errs.end() ? 2 : 3;
            r->err = it == errs.end() ? std::string() : it->second;
        }
        running_ -= batch.size();
        done_ += batch.size();
    }
    cv_done_.notify_all();
}

std::vector<Server::OutCopy> Server::run_batch(Lane &lane, std::vector<ServerRequest *> &batch) {
    Ctx &c = *lane.ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    c.set_current();
    const double t0 = now_s();
    for (const ServerRequest *r : batch)
        LC_REQUIRE(!(r->flags & LC_REQ_TEST_FAIL), "request marked LC_REQ_TEST_FAIL");
    // requests on the same geometry next to each other (stable: arrival order
    // within a geometry), so their sinograms and solution slots are contiguous
    // and one back-projection serves them all; a problem's result does not
    // depend on its place in the batch
    std::stable_sort(batch.begin(), batch.end(), [](const ServerRequest *a, const ServerRequest *b) {
        const void *ga = a->g3 ? (const void *)a->g3 : (const void *)a->g2;
        const void *gb = b->g3 ? (const void *)b->g3 : (const void *)b->g2;
        return std::less<const void *>()(ga, gb);
    });
    std::vector<ProbSpec> ps;
    for (ServerRequest *r : batch)
        ps.push_back(r->g3 ? ProbSpec{nullptr, (int)r->g3->n1, r->g3} : ProbSpec{r->g2, r->ns});
    // one stream per lane: the lanes already fill the hardware queues
    Batch2D b(c, ps, false, false);
    const double t1 = now_s();
    // inputs: right-hand sides (or sinograms back-projected into y), initial
    // guesses.  Every host array of the batch is packed into the lane's pinned
    // staging buffer and uploaded with one asynchronous copy per array into
    // one device staging region, so uploads and back-projections queue on the
    // stream without host waits.
    auto pad = [](size_t n) { return (n + 63) & ~(size_t)63; };
    size_t in_total = 0;
    bool any_x0 = false;
    for (ServerRequest *r : batch) {
        in_total += pad(r->in_n) + (r->x0p ? pad(r->x0_n) : 0);
        if (r->x0p) any_x0 = true;
    }
    float *hst = static_cast<float *>(lane.stage.get(sizeof(float) * std::max<size_t>(in_total, 1)));
    DBuf dst_in;
    dst_in.alloc(sizeof(float) * std::max<size_t>(in_total, 1));
    float *dst = dst_in.as<float>();
    size_t off = 0;
    // a run of 2-D sinograms on one geometry, contiguous in staging and in y
    const Geom2D *run_g = nullptr;
    float *run_s = nullptr, *run_y = nullptr;
    int run_ns = 0;
    size_t run_end = 0;          // staging offset just past the run
    int64_t run_yend = 0;        // y offset just past the run
    bool run_tight = false;      // the run's last sinogram has no staging padding
    auto flush = [&]() {
        if (run_ns) backproj2d_dev(c, *const_cast<Geom2D *>(run_g), run_s, run_y, run_ns);
        run_g = nullptr;
        run_ns = 0;
    };
    for (size_t i = 0; i < batch.size(); ++i) {
        ServerRequest *r = batch[i];
        const size_t src_n = r->in_n;
        std::memcpy(hst + off, r->in, sizeof(float) * src_n);
        float *s = dst + off;
        copy_in(c, s, hst + off, sizeof(float) * src_n, false);
        const size_t soff = off;
        off += pad(src_n);
        float *y = b.Y() + b.offset((int)i);
        if (r->has_sino && r->g2) {
            // one back-projection's fine-grid scratch stays near 1 GiB
            // (about 32 N^2 bytes per slice)
            const int64_t nn = r->g2->N * r->g2->N;
            const int run_cap = (int)std::max<int64_t>(1, ((int64_t)1 << 30) / (32 * nn));
            const bool extends = run_ns && run_g == r->g2 && run_end == soff && run_tight &&
                                 run_yend == b.offset((int)i) && run_ns + r->ns <= run_cap;
            if (!extends) {
                flush();
                run_g = r->g2;
                run_s = s;
                run_y = y;
            }
            run_ns += r->ns;
            run_end = off;
            run_tight = pad(src_n) == src_n;
            run_yend = b.offset((int)i) + b.size((int)i);
        } else {
            flush();
            if (r->has_sino) backproj3d_dev(c, *r->g3, s, y, nullptr);
            else LC_HIP(hipMemcpyAsync(y, s, sizeof(float) * src_n, hipMemcpyDeviceToDevice,
                                       c.stream()));
        }
        float *x = b.X() + b.offset((int)i);
        if (r->x0p) {
            std::memcpy(hst + off, r->x0p, sizeof(float) * r->x0_n);
            copy_in(c, dst + off, hst + off, sizeof(float) * r->x0_n, false);
            LC_HIP(hipMemcpyAsync(x, dst + off, sizeof(float) * r->x0_n,
                                  hipMemcpyDeviceToDevice, c.stream()));
            off += pad(r->x0_n);
        } else if (any_x0) {
            // (with no initial guess in the batch the solver clears x itself)
            LC_HIP(hipMemsetAsync(x, 0, sizeof(float) * (size_t)b.size((int)i), c.stream()));
        }
        // an initial guess staged after the sinogram breaks staging contiguity
        if (r->x0p) flush();
    }
    flush();
    c.sync();
    const double t2 = now_s();
    const ServerRequest &h = *batch.front();
    LC_HIP(hipEventRecord(lane.ev0, c.stream()));
    if (h.solver == 0) {
        std::vector<int> mi;
        for (ServerRequest *r : batch) mi.push_back(r->max_iter);
        b.cg(mi, h.tol, h.xtol, true, false, !any_x0);
    } else {
        std::vector<float> la, mu;
        for (ServerRequest *r : batch) {
            la.push_back(r->lambda);
            mu.push_back(r->mu);
        }
        b.split_bregman(la, mu, h.outer, h.inner, h.tol, h.xtol, (h.flags & LC_TV_ZERO_PAD) ? 1 : 0,
                        (h.flags & LC_TV_ANISO) != 0, true, false, !any_x0);
    }
    LC_HIP(hipEventRecord(lane.ev1, c.stream()));
    c.sync();
    const double t3 = now_s();
    BatchRecord rec;
    rec.lane = lane.index;
    rec.nreq = (int)batch.size();
    float d0 = 0.0f, d1 = 0.0f;
    LC_HIP(hipEventElapsedTime(&d0, epoch_, lane.ev0));
    LC_HIP(hipEventElapsedTime(&d1, epoch_, lane.ev1));
    rec.d0 = d0;
    rec.d1 = d1;
    const auto cs = b.cg_state();
    const auto ss = h.solver == 1 ? b.sb_state() : std::vector<SbState>();
    // results: the batch's whole solution buffer in one copy into the result
    // staging (problems are contiguous), once the previous batch's copy-out
    // has finished reading it; the per-request host copies are returned to
    // the lane loop, which runs them beside the lane's next batch
    std::vector<OutCopy> outs;
    if (lane.out_job.valid()) lane.out_job.get();
    float *hx = static_cast<float *>(lane.ostage.get(sizeof(float) * (size_t)b.total()));
    copy_out(c, hx, b.X(), sizeof(float) * (size_t)b.total(), false);
    c.sync();
    for (size_t i = 0; i < batch.size(); ++i) {
        ServerRequest *r = batch[i];
        outs.push_back(OutCopy{r, hx + b.offset((int)i), (size_t)b.size((int)i)});
        r->rep.iters = cs[i].iters;
        r->rep.outer = h.solver == 1 ? ss[i].outer : 0;
        r->rep.res = cs[i].res;
        r->rep.dx = h.solver == 1 ? ss[i].norm_diff : cs[i].dx;
        r->rep.stop = cs[i].stop;
        r->rep.pad = 0;
    }
    c.sync();
    const double t4 = now_s();
    rec.h0 = t0;
    rec.h1 = t4;
    std::lock_guard<std::mutex> lk2(m_);
    records_.push_back(rec);
    if (records_.size() > 65536) records_.pop_front();
    t_build_ += t1 - t0;
    t_input_ += t2 - t1;
    t_solve_ += t3 - t2;
    t_output_ += t4 - t3;
    return outs;
}

} // namespace lc

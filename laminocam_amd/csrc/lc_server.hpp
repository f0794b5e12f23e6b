// Reconstruction request server (continuous batching, HBM-budget admission).
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <future>
#include <thread>
#include <vector>

#include "lc_engine.hpp"

namespace lc {

struct ServerRequest {
    uint64_t id = 0;
    Geom2D *g2 = nullptr;      // 2-D stack of ns slices, or
    Geom3D *g3 = nullptr;      // one 3-D volume
    int ns = 1;
    int solver = 0;            // 0 CG, 1 Split-Bregman
    int max_iter = 50;         // CG iterations
    int outer = 1, inner = 1;  // Split-Bregman iterations
    double tol = 1e-5, xtol = 1e-5;
    float lambda = 0.1f, mu = 10.0f;
    uint32_t flags = 0;        // LC_TV_ZERO_PAD, LC_TV_ANISO
    bool has_sino = false;
    std::vector<float> sino;   // sinogram / projections (back-projected on the device)
    std::vector<float> rhs;    // right-hand side when no sinogram
    std::vector<float> x0;     // empty: start from zero
    // the input the lane reads (sinogram or right-hand side) and the initial
    // guess (x0p == nullptr: zero): into the vectors above, or the caller's
    // arrays when borrowed
    const float *in = nullptr;
    size_t in_n = 0;
    const float *x0p = nullptr;
    size_t x0_n = 0;
    // results: into the caller's x_out when given, else into x
    float *x_out = nullptr;
    std::vector<float> x;
    lc_report rep{};
    size_t bytes = 0;          // device bytes held while running
    double t_submit = 0, t_start = 0, t_done = 0;
    int state = 0;             // 0 queued, 1 running, 2 done, 3 failed
    int bypassed = 0;          // batches admitted past this request while it headed the queue
    std::string err;
};

// one batch a lane ran: lane, requests, host start / end (seconds, steady
// clock), device start / end of the solve (ms since the server started)
struct BatchRecord {
    int lane = 0, nreq = 0;
    double h0 = 0, h1 = 0, d0 = 0, d1 = 0;
};

size_t request_bytes(const ServerRequest &r);

class Server {
  public:
    Server(Ctx &main, size_t budget, int lanes, int max_batch);
    ~Server();
    uint64_t submit(std::unique_ptr<ServerRequest> r);
    bool poll(uint64_t id);
    bool wait(uint64_t id, double timeout_s);   // < 0: no timeout
    std::unique_ptr<ServerRequest> take(uint64_t id);
    // in-flight bytes, queued, running, finished, batches, budget, then the
    // summed lane seconds spent building batches, loading inputs (incl.
    // back-projection), solving, and reading results back
    void stats(double *out);
    // the most recent batch records, oldest first (at most max_rows)
    std::vector<BatchRecord> timeline(size_t max_rows);

  private:
    // one request's solution, from the lane's result staging to its host array
    struct OutCopy {
        ServerRequest *r;
        const float *src;
        size_t n;
    };
    struct Lane {
        int index = 0;
        MemPool pool;                // outlives the context's buffers
        PinnedBuf stage;             // host staging of the lane's inputs
        PinnedBuf ostage;            // host staging of the lane's results
        std::future<void> out_job;   // the previous batch's copy-out (reads ostage)
        std::unique_ptr<Ctx> ctx;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;   // device start / end of a batch's solve
        std::thread th;
    };
    void lane_loop(Lane &lane);
    std::vector<OutCopy> run_batch(Lane &lane, std::vector<ServerRequest *> &batch);
    // host copies of a batch's results; a request whose solution is not
    // finite fails with that message (errs[i], empty for success)
    static void deliver(const std::vector<OutCopy> &outs, std::map<ServerRequest *, std::string> &errs);
    void finish(const std::vector<ServerRequest *> &batch, const std::map<ServerRequest *, std::string> &errs);

    Ctx &main_;
    size_t budget_;
    int max_batch_;
    std::vector<std::unique_ptr<Lane>> lanes_;
    std::mutex m_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<uint64_t> queue_;
    std::map<uint64_t, std::unique_ptr<ServerRequest>> reqs_;
    uint64_t next_id_ = 0;
    size_t inflight_ = 0, queued_ = 0, running_ = 0, done_ = 0, batches_ = 0;
    double t_build_ = 0, t_input_ = 0, t_solve_ = 0, t_output_ = 0;
    bool stop_ = false;
    hipEvent_t epoch_ = nullptr;        // recorded when the server starts
    std::deque<BatchRecord> records_;   // the most recent batches (guarded by m_)
};

} // namespace lc

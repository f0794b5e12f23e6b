// Device context, buffers and twiddle cache.
#pragma once

#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "lc_common.hpp"
#include "lc_fft.hpp"

namespace lc {

// One process-wide lock taken around every stream capture (begin capture to
// graph instantiation) and around every call that allocates or frees device
// or pinned host memory.  hipFree and hipHostFree synchronise the whole
// device, and a capture on one thread's stream must not race with a
// device-wide synchronisation or an allocation from another thread (serving
// lanes capture concurrently, a submitting thread builds geometries beside
// them).  Recursive: an allocation inside a held section takes it again.
inline std::recursive_mutex &capture_mutex() {
    static std::recursive_mutex m;
    return m;
}

// Cache of device blocks for one worker thread.  hipFree synchronises the
// whole device, so a thread whose buffers come and go (a serving lane, one
// batch after another) keeps them here and reuses them; blocks are sized in
// bins of 1, 1.25, 1.5, 1.75 x 2^k (at least 64 KiB) and freed only when the
// pool is destroyed.
class MemPool {
  public:
    MemPool() = default;
    MemPool(const MemPool &) = delete;
    MemPool &operator=(const MemPool &) = delete;
    ~MemPool() {
        std::lock_guard<std::recursive_mutex> cl(capture_mutex());
        for (auto &kv : free_) (void)hipFree(kv.second);
    }
    static size_t bin(size_t bytes) {
        size_t b = 64 << 10;
        while (b < bytes) b <<= 1;
        if (b <= (64 << 10)) return b;
        const size_t q = b >> 3;   // candidates 4q, 5q, 6q, 7q, 8q
        for (size_t m = 4; m <= 8; ++m)
            if (m * q >= bytes) return m * q;
        return b;
    }
    void *get(size_t bytes, size_t &got) {
        got = bin(bytes);
        std::lock_guard<std::mutex> lk(m_);
        auto it = free_.find(got);
        if (it != free_.end()) {
            void *p = it->second;
            free_.erase(it);
            return p;
        }
        void *p = nullptr;
        std::lock_guard<std::recursive_mutex> cl(capture_mutex());
        hipError_t e = hipMalloc(&p, got);
        if (e != hipSuccess)
            throw Error(LC_ENOMEM, "hipMalloc of " + std::to_string(got) + " bytes failed: " +
                                       hipGetErrorString(e));
        return p;
    }
    void put(void *p, size_t got) {
        std::lock_guard<std::mutex> lk(m_);
        free_.emplace(got, p);
    }

  private:
    std::mutex m_;
    std::multimap<size_t, void *> free_;
};

// allocations made by this thread go through this pool when set
inline thread_local MemPool *tl_pool = nullptr;

// A fixed device region handed out front to back with no hipMalloc: while
// tl_arena is set, the DBuf allocations of this thread come from it (256-byte
// aligned) and are never freed.  A request beyond its end fails with
// LC_EINVAL.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    const char *who = "";
    void *take(size_t bytes) {
        const size_t off = (used + 255) & ~(size_t)255;
        if (off > cap || bytes > cap - off)
            throw Error(LC_EINVAL, std::string(who) + ": the problem needs more than the " +
                                       std::to_string(cap >> 20) + " MiB device scratch");
        used = off + bytes;
        return base + off;
    }
};
inline thread_local Arena *tl_arena = nullptr;

// Grow-only pinned host buffer (hipHostMalloc): staging for asynchronous
// host <-> device copies.  Grows by at least 1.5x so that a sequence of
// batches of similar size reallocates rarely.
class PinnedBuf {
  public:
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() {
        std::lock_guard<std::recursive_mutex> cl(capture_mutex());
        if (p_) (void)hipHostFree(p_);
    }
    // a buffer of at least `bytes`; the caller has no copy in flight on it
    void *get(size_t bytes) {
        if (bytes > n_) {
            std::lock_guard<std::recursive_mutex> cl(capture_mutex());
            const size_t want = std::max(bytes, n_ + n_ / 2);
            if (p_) (void)hipHostFree(p_);
            p_ = nullptr;
            n_ = 0;
            hipError_t e = hipHostMalloc(&p_, want, hipHostMallocDefault);
            if (e != hipSuccess) {
                p_ = nullptr;
                throw Error(LC_ENOMEM, "hipHostMalloc of " + std::to_string(want) +
                                           " bytes failed: " + hipGetErrorString(e));
            }
            n_ = want;
        }
        return p_;
    }

  private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

// RAII device allocation
class DBuf {
  public:
    DBuf() = default;
    explicit DBuf(size_t bytes) { alloc(bytes); }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept
        : p_(o.p_), n_(o.n_), cap_(o.cap_), pool_(o.pool_), arena_(o.arena_) {
        o.p_ = nullptr; o.n_ = 0; o.cap_ = 0; o.pool_ = nullptr; o.arena_ = false;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_; cap_ = o.cap_; pool_ = o.pool_; arena_ = o.arena_;
            o.p_ = nullptr; o.n_ = 0; o.cap_ = 0; o.pool_ = nullptr; o.arena_ = false;
        }
        return *this;
    }
    void alloc(size_t bytes) {
        release();
        if (bytes == 0) return;
        if (tl_arena) {
            p_ = tl_arena->take(bytes);
            n_ = cap_ = bytes;
            arena_ = true;
            return;
        }
        if (tl_pool) {
            p_ = tl_pool->get(bytes, cap_);
            pool_ = tl_pool;
            n_ = bytes;
            return;
        }
        std::lock_guard<std::recursive_mutex> cl(capture_mutex());
        hipError_t e = hipMalloc(&p_, bytes);
        if (e != hipSuccess) {
            p_ = nullptr;
            throw Error(LC_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                       hipGetErrorString(e));
        }
        n_ = cap_ = bytes;
    }
    // grow-only reallocation (contents not preserved)
    void reserve(size_t bytes) {
        if (bytes > n_) alloc(bytes);
    }
    void release() {
        if (p_ && !arena_) {
            if (pool_) {
                pool_->put(p_, cap_);
            } else {
                std::lock_guard<std::recursive_mutex> cl(capture_mutex());
                (void)hipFree(p_);
            }
        }
        p_ = nullptr;
        n_ = cap_ = 0;
        pool_ = nullptr;
        arena_ = false;
    }
    template <typename T = void> T *as() const { return reinterpret_cast<T *>(p_); }
    size_t bytes() const { return n_; }

  private:
    void *p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
    MemPool *pool_ = nullptr;
    bool arena_ = false;   // carved from tl_arena: nothing to free
};

// Block shapes of the three Toeplitz passes for one image size:
// row = P1/P3 shape index, col = P2 shape index, B = P2 columns per block.
struct ToepVariant {
    int row = 0, col = 0, B = 0;
};

class Ctx {
  public:
    // per image size N: kernel shapes chosen by the autotuner (or installed)
    std::map<int, ToepVariant> toep_tuning;
    // per 3-D fine-grid shape (nf0, nf1, nf2): the type-1 spread's bin shape
    // (an index into the compiled shapes; 0 when none was chosen)
    std::map<std::array<int, 3>, int> nu3_tuning;
    int nu3_shape(const int nf[3]) const {
        auto it = nu3_tuning.find({nf[0], nf[1], nf[2]});
        return it == nu3_tuning.end() ? 0 : it->second;
    }
    // per 2-D fine-grid size nf: the type-1 spread's bin shape
    std::map<int, int> nu2_tuning;
    int nu2_shape(int nf) const {
        auto it = nu2_tuning.find(nf);
        return it == nu2_tuning.end() ? 0 : it->second;
    }
    explicit Ctx(int device);
    ~Ctx();
    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    void sync();
    void set_current() const;

    // FP64-generated twiddle table of length L (cached per context)
    template <typename T> const cpx<T> *twiddles(int L);

    // scratch buffers owned by the context (grow only).  Valid until the next
    // request for the same slot; callers on one context are serialised.
    void *scratch(int slot, size_t bytes);

    std::mutex mu;   // serialises API calls that share this context

    // Auxiliary streams (default priority) on which independent work of one
    // call may run beside the main stream: fork(n) makes the first n wait for
    // the main stream's work so far, join(n) makes the main stream wait for
    // theirs.  Inside a stream capture these become graph dependencies.
    // They are created on first request (ensure_aux, outside any capture):
    // a process gets few hardware queues (four by default) and streams beyond
    // them share one, so a context that never forks holds one stream only.
    static constexpr int kAux = 3;
    void ensure_aux();
    hipStream_t aux(int i) const { return aux_[i]; }
    void fork(int n);
    void join(int n);
    // the stream launches go to (the main stream unless a StreamScope is active)
    class StreamScope {
      public:
        StreamScope(Ctx &c, hipStream_t s) : c_(c), prev_(c.stream_) { c.stream_ = s; }
        ~StreamScope() { c_.stream_ = prev_; }

      private:
        Ctx &c_;
        hipStream_t prev_;
    };
    hipStream_t main_stream() const { return main_; }

  private:
    int device_;
    hipStream_t stream_ = nullptr, main_ = nullptr;
    hipStream_t aux_[kAux] = {};
    hipEvent_t fork_ev_ = nullptr, join_ev_[kAux] = {};
    std::map<std::pair<int, int>, DBuf> tw_;
    std::vector<DBuf> scratch_;
};

template <typename T> const cpx<T> *Ctx::twiddles(int L) {
    auto key = std::make_pair(L, (int)sizeof(T));
    auto it = tw_.find(key);
    if (it != tw_.end()) return it->second.as<cpx<T>>();
    // a cached table must not live in an arena, whose memory the next user overwrites
    LC_REQUIRE(!tl_arena, "no FFT twiddles of length " + std::to_string(L) +
                              " prepared on this context (create the geometry on it without "
                              "LC_GEOM_LAZY_PSF)");
    DBuf b(sizeof(cpx<T>) * (size_t)L);
    hipLaunchKernelGGL(k_twiddles<T>, dim3(ceil_div(L, 256)), dim3(256), 0, stream_,
                       b.as<cpx<T>>(), L);
    LC_HIP(hipGetLastError());
    const cpx<T> *p = b.as<cpx<T>>();
    tw_.emplace(key, std::move(b));
    return p;
}

// An executable graph captured from the launches `body` issues on a
// context's stream.  Capture and instantiation run under capture_mutex().
// The destructor first waits for the stream: an executable graph is not
// destroyed while launches of it are still queued or running.
class CapturedGraph {
  public:
    template <typename F> CapturedGraph(Ctx &c, F &&body) : c_(c) {
        std::lock_guard<std::recursive_mutex> cl(capture_mutex());
        LC_HIP(hipStreamBeginCapture(c.stream(), hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t partial = nullptr;
            if (hipStreamEndCapture(c.stream(), &partial) == hipSuccess && partial)
                (void)hipGraphDestroy(partial);
            throw;
        }
        LC_HIP(hipStreamEndCapture(c.stream(), &g_));
        const hipError_t e = hipGraphInstantiate(&ge_, g_, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g_);
            g_ = nullptr;
            LC_HIP(e);
        }
    }
    CapturedGraph(const CapturedGraph &) = delete;
    CapturedGraph &operator=(const CapturedGraph &) = delete;
    ~CapturedGraph() {
        (void)hipStreamSynchronize(c_.stream());
        std::lock_guard<std::recursive_mutex> cl(capture_mutex());
        if (ge_) (void)hipGraphExecDestroy(ge_);
        if (g_) (void)hipGraphDestroy(g_);
    }
    void launch() { LC_HIP(hipGraphLaunch(ge_, c_.stream())); }

  private:
    Ctx &c_;
    hipGraph_t g_ = nullptr;
    hipGraphExec_t ge_ = nullptr;
};

// copy helpers honouring a "pointers are on the device" flag
void copy_in(Ctx &c, void *dst, const void *src, size_t bytes, bool src_dev);
void copy_out(Ctx &c, void *dst, const void *src, size_t bytes, bool dst_dev);

} // namespace lc

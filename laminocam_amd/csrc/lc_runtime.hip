// Device context implementation.
#include "lc_runtime.hpp"

namespace lc {

Ctx::Ctx(int device) : device_(device) {
    int n = 0;
    LC_HIP(hipGetDeviceCount(&n));
    LC_REQUIRE(device >= 0 && device < n, "lc_ctx_create: device index out of range");
    LC_HIP(hipSetDevice(device));
    // default-priority, non-blocking stream
    std::lock_guard<std::recursive_mutex> cl(capture_mutex());
    LC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    main_ = stream_;
}

void Ctx::ensure_aux() {
    if (fork_ev_) return;
    std::lock_guard<std::recursive_mutex> cl(capture_mutex());
    for (int i = 0; i < kAux; ++i) {
        LC_HIP(hipStreamCreateWithFlags(&aux_[i], hipStreamNonBlocking));
        LC_HIP(hipEventCreateWithFlags(&join_ev_[i], hipEventDisableTiming));
    }
    LC_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
}

void Ctx::fork(int n) {
    LC_REQUIRE(fork_ev_ != nullptr && n <= kAux, "context: fork before ensure_aux");
    LC_HIP(hipEventRecord(fork_ev_, stream_));
    for (int i = 0; i < n; ++i) LC_HIP(hipStreamWaitEvent(aux_[i], fork_ev_, 0));
}

void Ctx::join(int n) {
    for (int i = 0; i < n; ++i) {
        LC_HIP(hipEventRecord(join_ev_[i], aux_[i]));
        LC_HIP(hipStreamWaitEvent(stream_, join_ev_[i], 0));
    }
}

Ctx::~Ctx() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (int i = 0; i < kAux; ++i)
        if (aux_[i]) (void)hipStreamSynchronize(aux_[i]);
    tw_.clear();
    scratch_.clear();
    std::lock_guard<std::recursive_mutex> cl(capture_mutex());
    for (int i = 0; i < kAux; ++i) {
        if (join_ev_[i]) (void)hipEventDestroy(join_ev_[i]);
        if (aux_[i]) (void)hipStreamDestroy(aux_[i]);
    }
    if (fork_ev_) (void)hipEventDestroy(fork_ev_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Ctx::set_current() const { LC_HIP(hipSetDevice(device_)); }

void Ctx::sync() { LC_HIP(hipStreamSynchronize(stream_)); }

void *Ctx::scratch(int slot, size_t bytes) {
    if ((int)scratch_.size() <= slot) scratch_.resize(slot + 1);
    if (scratch_[slot].bytes() < bytes) {
        LC_REQUIRE(!tl_arena, "context scratch cannot grow inside an arena");
        LC_HIP(hipStreamSynchronize(stream_));
        scratch_[slot].alloc(bytes);
    }
    return scratch_[slot].as();
}

void copy_in(Ctx &c, void *dst, const void *src, size_t bytes, bool src_dev) {
    if (bytes == 0) return;
    LC_HIP(hipMemcpyAsync(dst, src, bytes, src_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          c.stream()));
}

void copy_out(Ctx &c, void *dst, const void *src, size_t bytes, bool dst_dev) {
    if (bytes == 0) return;
    LC_HIP(hipMemcpyAsync(dst, src, bytes, dst_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                          c.stream()));
}

} // namespace lc

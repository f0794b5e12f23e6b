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
}

Ctx::~Ctx() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    tw_.clear();
    scratch_.clear();
    std::lock_guard<std::recursive_mutex> cl(capture_mutex());
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

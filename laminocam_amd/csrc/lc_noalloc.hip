// The allocation-free 2-D CG solve (lc_cg2d_noalloc): lc_cg2d whose device
// buffers come from a fixed array of the library instead of hipMalloc, for
// callers that cannot allocate device memory.
//
// The array below has one copy per GPU the library runs on and is held until
// the process ends.  Host code never names it as an address: its device
// address on each GPU is looked up once with hipGetSymbolAddress when the
// first context on that GPU is created, and only that address is used.
// Calls on one GPU share the array; a per-GPU lock makes them run one after
// another, and every call overwrites what the previous one left there.
#include <mutex>
#include <string>

#include "lc_engine.hpp"

static constexpr size_t kScratchBytes = (size_t)LC_NOALLOC_SCRATCH_MIB << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace lc {

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_setup_mu;
void *g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch`, per device
std::mutex g_call_mu[kMaxDevices];       // one call at a time per device

} // namespace

void noalloc_setup(int device) {
    LC_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (g_scratch_ptr[device]) return;
    int cur = -1;
    LC_HIP(hipGetDevice(&cur));
    LC_REQUIRE(cur == device, "noalloc_setup: the device must be current");
    void *ptr = nullptr;
    LC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    LC_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = ptr;
}

CgState cg2d_noalloc(Ctx &c, Geom2D &g, const float *rhs, float *x, int ns, int64_t max_iter,
                     double tol, double xtol, uint32_t flags) {
    const char *who = "lc_cg2d_noalloc";
    LC_REQUIRE(flags & LC_DEVICE_PTRS, std::string(who) + ": rhs and x must be device pointers "
                                                          "(LC_DEVICE_PTRS)");
    LC_REQUIRE(!(flags & LC_GRAPH), std::string(who) + ": LC_GRAPH is not supported");
    LC_REQUIRE(g.psf_ready, std::string(who) + ": the geometry's PSF is not built (a geometry "
                                               "made with LC_GEOM_LAZY_PSF builds it on its "
                                               "first Toeplitz apply)");
    LC_REQUIRE(max_iter >= 0 && max_iter <= (1 << 30), std::string(who) + ": bad max_iter");
    const int dev = c.device();
    LC_REQUIRE(dev >= 0 && dev < kMaxDevices, std::string(who) + ": device index out of range");
    std::lock_guard<std::mutex> call(g_call_mu[dev]);
    std::lock_guard<std::mutex> lk(c.mu);
    c.set_current();
    if (!g_scratch_ptr[dev]) noalloc_setup(dev);   // (contexts set it up when created)
    Arena arena;
    arena.base = static_cast<char *>(g_scratch_ptr[dev]);
    arena.cap = kScratchBytes;
    arena.who = who;
    struct Use {
        Arena *prev;
        explicit Use(Arena *a) : prev(tl_arena) { tl_arena = a; }
        ~Use() { tl_arena = prev; }
    } use(&arena);
    // the batch's vectors, Toeplitz intermediate, tables and partial sums are
    // all carved from the array; a problem that does not fit fails here,
    // before anything is launched on the caller's data
    Batch2D b(c, {ProbSpec{&g, ns}});
    const size_t bytes = sizeof(float) * (size_t)b.size(0);
    const bool x0_zero = (flags & LC_X0_ZERO) != 0;
    copy_in(c, b.Yv(), rhs, bytes, true);
    if (!x0_zero) copy_in(c, b.Xv(), x, bytes, true);
    // no graph, no host read of the scalars between iterations: finished
    // problems' blocks exit on the device-side active flag
    b.cg({(int)max_iter}, tol, xtol, false, false, x0_zero);
    copy_out(c, x, b.Xv(), bytes, true);
    c.sync();
    return b.cg_state()[0];
}

} // namespace lc

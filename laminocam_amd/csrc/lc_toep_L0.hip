// Toeplitz pass kernels instantiated for FFT length 0 (0: any length, runtime plan).
#include "lc_engine.hpp"

namespace lc {
extern __shared__ __align__(16) char lc_smem[];
#define LC_TOEP_LF 0
#include "lc_toep_kernels.inc"
} // namespace lc

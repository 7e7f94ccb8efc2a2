// complex128 instantiation of the tile-pass kernels (T = 12):
// forward 16 amplitudes x 256 threads, backward 8 amplitudes x 512 threads.
#include "sv_impl.h"

namespace jq {
hipError_t launch_sv_f64(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  if (T != 12) return hipErrorInvalidValue;
  switch (which) {
    case 0: return launch_forward<double, 4, 256, false>(a, ntiles, batch, smem, st);
    case 1: return launch_backward<double, 3, 512, false>(a, ntiles, batch, smem, st);
    case 2: return launch_forward<double, 4, 256, true>(a, ntiles, batch, smem, st);
    case 3: return launch_backward<double, 3, 512, true>(a, ntiles, batch, smem, st);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace jq

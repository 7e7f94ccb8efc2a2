// complex64 instantiation of the tile-pass kernels (T = 13):
// forward 32 amplitudes x 256 threads, backward 16 amplitudes x 512 threads.
#include "sv_impl.h"

namespace jq {
hipError_t launch_sv_f32(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  if (T != 13) return hipErrorInvalidValue;
  switch (which) {
    case 0: return launch_forward<float, 5, 256, false>(a, ntiles, batch, smem, st);
    case 1: return launch_backward<float, 4, 512, false>(a, ntiles, batch, smem, st);
    case 2: return launch_forward<float, 5, 256, true>(a, ntiles, batch, smem, st);
    case 3: return launch_backward<float, 4, 512, true>(a, ntiles, batch, smem, st);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace jq

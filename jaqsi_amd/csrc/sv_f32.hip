// complex64 instantiation of the tile-pass kernels (T = 13: 256 threads x 32 amplitudes).
#include "sv_impl.h"

namespace jq {
hipError_t launch_sv_f32(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  if (T == 13) return launch_cfg<float, 5, 256>(which, a, ntiles, batch, smem, st);
  return hipErrorInvalidValue;
}
}  // namespace jq

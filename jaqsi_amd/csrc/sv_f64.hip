// complex128 instantiation of the tile-pass kernels (T = 12: 256 threads x 16 amplitudes).
#include "sv_impl.h"

namespace jq {
hipError_t launch_sv_f64(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  if (T == 12) return launch_cfg<double, 4, 256>(which, a, ntiles, batch, smem, st);
  return hipErrorInvalidValue;
}
}  // namespace jq

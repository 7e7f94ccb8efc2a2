// complex64 instantiation of the tile-pass kernels (T = 13):
// forward 16 amplitudes x 512 threads (<= 128 VGPRs: two workgroups, 16
// waves per CU), backward 8 amplitudes x 1024 threads (psi and lambda, 16
// waves per CU).  Plans with 4x4 or dense 3-wire blocks run their backward
// with 16 amplitudes x 512 threads instead (256-VGPR budget): under the
// 1024-thread cap those blocks would push the state arrays into scratch.
#include "sv_impl.h"

namespace jq {
hipError_t launch_sv_f32(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  if (T != 13) return hipErrorInvalidValue;
  switch (which) {
    case 0: return launch_forward<float, 4, 512, false>(a, ntiles, batch, smem, st);
    case 1: return launch_backward<float, 3, 1024, false>(a, ntiles, batch, smem, st);
    case 2: return launch_forward<float, 4, 512, true>(a, ntiles, batch, smem, st);
    // 3 (full variant, narrow backward) is never used: full plans take the wide backward
    case 5: return launch_backward<float, 4, 512, false>(a, ntiles, batch, smem, st);
    case 7: return launch_backward<float, 4, 512, true>(a, ntiles, batch, smem, st);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace jq

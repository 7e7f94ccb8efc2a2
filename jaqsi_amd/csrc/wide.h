// Launch interface of the wide-block pass kernels (wide.hip).
//
// A wide block (program.h: AK_WIDE) is one dense matrix on up to 5 wires,
// or one diagonal on up to 6, applied to the whole vector in a pass of its
// own.  Its adjoint reads the checkpointed input state of the pass, so the
// block is never inverted: this is how singular or ill-conditioned channel
// superoperators, dense gates of three or more target wires and wide
// diagonals get their cotangents.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sv_kernels.h"

namespace jq {

struct WideGeom {
  int k;            // wires
  int D;            // 2^k
  int diag;         // diagonal storage (D entries) instead of D x D
  int gb[6];        // global bit of wire j (matrix-index bit k-1-j)
  int sorted[6];    // the same bits ascending
  int64_t rpw;      // rest indices (all wire bits zero) per workgroup
  int lgS;          // log2 of the rest indices per LDS sub-chunk
  int lgD;          // log2 D
  int64_t mat;      // pool offset of the block matrix
  int slot;         // first cotangent real within the pass (-1: no cotangent)
  int slot_base;    // the pass's partial area (reals)
  int ntiles;       // workgroups per circuit = the plan's tile count
};

// Fill the geometry of a wide gate record `g` (program.h words) for an
// n-wire vector split into `ntiles` workgroups per circuit.
void wide_geom(const int32_t* g, int n_wires, int ntiles, int slot_base, int dtype, WideGeom& out);

// dir 0: forward (a.src -> a.dst, in place allowed; a.src null = |0..0>);
// dir 1: backward (a.src = the pass's input state, a.lam updated in place,
// cotangent partials into a.grad_part).  Grid: ntiles x batch.
hipError_t launch_wide(int dir, int dtype, const SvArgs& a, const WideGeom& g, int batch, hipStream_t st);

}  // namespace jq

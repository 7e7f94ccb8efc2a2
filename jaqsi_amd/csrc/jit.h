// Plan-specialised tile-pass kernels, compiled at run time with hipRTC.
//
// The generic kernels interpret the program words of a plan: per gate they
// decode a record and dispatch on its kind and register bits.  For a fixed
// plan the same sequence can be written out as straight-line code with the
// register bits, control masks, pool offsets and layouts as literals; the
// compiler then keeps the state in registers across the whole pass, turns
// register permutations into renaming and drops every decode.  jit_source()
// writes that code (a forward and a backward kernel per distinct pass) over the
// device building blocks of sv_impl.h, and jit_compile() turns it into a
// gfx950 code object.  Both run on the host only: no HIP runtime call is
// made, so they may run on a background thread.
#pragma once
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

namespace jq {

struct JitSpec {
  const std::vector<int32_t>* words_f = nullptr;  // forward program
  const std::vector<int32_t>* words_b = nullptr;  // backward program
  int dtype = 0;                                  // 0 complex64, 1 complex128
  // Backward: persistent loop over (circuit, tile) items with next-item
  // prefetch (one resident workgroup per CU, so the loads would otherwise not
  // overlap any work).  A persistent forward measured slower: its prefetch
  // registers halve the resident workgroups.
  bool persist_bwd = true;
  // Forward (complex64): the tile streams into LDS and the pass starts in its
  // first substep's layout (sv_forward_dma, one tile per workgroup so two
  // workgroups fit per CU; a persistent variant with a separate prefetch
  // tile measured slower).
  bool dma = true;
};

// Unique pass bodies of a plan.  Passes whose gate sequences differ only by
// the matrix offsets share one kernel: each pass runs kernel kernel_of[dir][q]
// with SvArgs::pool_off = pool_off[dir][q].
struct JitPlan {
  struct Kernel {
    int dir;            // 0 forward, 1 backward
    std::string body;   // generated statements of the pass body (from the canonical layout)
    std::string first;  // layout words of the first substep visited, empty when canonical
    std::string body0;  // the statements after the initial re-layout
    std::string last;   // layout words at the end of the body, empty when canonical
    // HBM access under the first / last layout instead of the canonical one
    // (0: canonical, 1: register pairs of adjacent amplitudes, 2: single
    // amplitudes); `full`: n_wires >= T.
    int load = 0, store = 0;
    bool full = false;
  };
  std::vector<Kernel> kernels;
  std::vector<int> kernel_of[2];
  std::vector<int64_t> pool_off[2];
};

void jit_plan(const JitSpec& spec, JitPlan& out);

// Kernel names: jq_sv_forward_s<n>, jq_sv_backward_s<n> (n unique in the process).
std::string jit_kernel_name(int dir, int n);

// HIP C++ source of the listed kernels of `plan`; kernel ids[i] is emitted
// as the extern "C" kernel names[i].
std::string jit_source(const JitSpec& spec, const JitPlan& plan, const std::vector<int>& ids,
                       const std::vector<std::string>& names);

// Key under which a compiled kernel can serve any plan: everything the
// generated kernel depends on (direction, dtype, register tile, loop variant,
// body text).
std::string jit_key(const JitSpec& spec, const JitPlan& plan, int id);

// Compile `src` for gfx950.  Returns false and fills `log` on failure.
bool jit_compile(const std::string& src, std::vector<char>& code, std::string& log);

}  // namespace jq

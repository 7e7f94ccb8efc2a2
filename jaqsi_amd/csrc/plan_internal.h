// Planner output shared between plan.cpp and the C-ABI layer.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "jaqsi_amd.h"

namespace jq {

struct PlanError : std::runtime_error {
  int code;
  PlanError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct SlotMap {
  int pass;   // pass index
  int local;  // first real within the pass's per-tile slot area
  int bar;    // first real within the circuit's cotangent array
  int size;   // reals
  int base;   // the pass's slot_base (reals)
};

struct HostProgram {
  std::vector<int32_t> words;         // scheduled program for the forward sweep (program.h)
  std::vector<int32_t> words_b;       // the same passes, laid out for the backward sweep
  std::vector<int32_t> block_table;   // BW words per device block
  std::vector<int32_t> member_list;   // member indices, grouped by block
  std::vector<int32_t> member_table;  // MW words per member
  std::vector<double> member_values;  // constant angle per member
  std::vector<SlotMap> slot_map;      // grad slot -> cotangent reals
  int T = 0, RB = 0, n_wires = 0, ntiles_log = 0, ntiles = 1;
  int pool = 0;           // complex elements of block matrices per circuit
  int bars = 0;           // complex elements of cotangents per circuit
  int partial_reals = 0;  // reals of per-tile partial sums per circuit
  int n_dev_blocks = 0, n_passes = 0, n_substeps = 0;
};

// slot_budget: most cotangent reals one pass may flush (they live in LDS
// next to the tiles); a pass closes early rather than exceed it.
// max_blocks > 0 caps the blocks of a tile pass (a pass's straight-line
// specialised code grows with its gate count).
void build_program(const jq_plan_desc& d, int T, int RB, int RBb, HostProgram& out, int slot_budget,
                   int max_blocks = 0);

// Trailing blocks folded into the observables.  A unitary block that
// permutes basis states by an affine map over GF(2) (x -> A x ^ b: X, CX,
// SWAP and their products) or multiplies them by phases (diagonal blocks),
// and that nothing later in the circuit touches, does not change the
// expectation value of a diagonal observable measured after it except by
// moving which index bits it reads: the final bit of wire w is the parity of
// the pre-tail bits in rows[w], XOR bit w of flip.  Those blocks are then
// left out of the passes, and a Pauli-Z word on the final vector becomes a
// Pauli-Z word (with the sign of `flip`) on the vector the passes compute.
struct TailFold {
  int n_folded = 0;            // ops left out of the passes
  std::vector<uint64_t> rows;  // per wire: wire mask (bit v = wire v)
  uint64_t flip = 0;           // per wire
  bool operator==(const TailFold& o) const {
    return n_folded == o.n_folded && (n_folded == 0 || (rows == o.rows && flip == o.flip));
  }
};

// Splits d.ops into the ops the passes run (`kept`, in order) and the
// folded tail.  Only parameter-free ops fold (their angle gradients would be
// zero), never non-unitary ones or explicit matrices, and at least one op
// always stays.
void fold_tail(const jq_plan_desc& d, std::vector<jq_op>& kept, TailFold& f);

}  // namespace jq

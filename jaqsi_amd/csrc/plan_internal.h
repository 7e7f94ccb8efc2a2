// Planner output shared between plan.cpp and the C-ABI layer.
#pragma once
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

}  // namespace jq

// Host planner: lowered op list -> fused blocks -> scheduled device program.
//
// Fusion follows the reference's plan rule (jaqsi/simulation.py:214-240): an
// op multiplies into the previous block on the same wire tuple when no block
// in between touched any of those wires; structure tags combine as in
// _fused_structure (simulation.py:202-211).  The reference then composes
// runs of permutations/diagonals into whole-state gathers
// (simulation.py:243-335); here those runs are instead scheduled into tile
// passes, where permutations and diagonals cost no HBM traffic of their own.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "jaqsi_amd.h"
#include "plan_internal.h"
#include "program.h"

namespace jq {

namespace {

struct HB {  // host block
  std::vector<int> wires;
  int structure = 0;
  int controls = 0;
  std::vector<int> perm;
  std::vector<int> members;
  int kind = AK_NOP;
  int k = 0, D = 0, diagf = 0, offs = 0;
  bool grad = false;
  bool nonunitary = false;
  bool quat = false;  // MC1 block of the form [[a, -b*], [b, a*]] for every angle
  bool wide = false;  // applied in a pass of its own (AK_WIDE)
  double cond = 1.0;  // bound on the condition number: product over its non-unitary ops
  int mat = -1, inv = -1, bar = -1, bar_size = 0;
  std::vector<int> targets;  // wires that must sit in registers, in matrix order
  std::vector<int> ctrls;    // control wires
};

bool is_identity(const std::vector<int>& p) {
  for (size_t i = 0; i < p.size(); ++i)
    if (p[i] != (int)i) return false;
  return true;
}

bool is_mcx(const std::vector<int>& p) {  // X on the last wire, all others controls
  int d = (int)p.size();
  for (int i = 0; i < d - 2; ++i)
    if (p[i] != i) return false;
  return p[d - 2] == d - 1 && p[d - 1] == d - 2;
}

bool is_mcswap(const std::vector<int>& p) {  // swap of the last two wires under controls
  int d = (int)p.size();
  if (d < 4) return false;
  int s = d - 4;
  for (int i = 0; i < d; ++i) {
    int want = i;
    if (i == s + 1) want = s + 2;
    if (i == s + 2) want = s + 1;
    if (p[i] != want) return false;
  }
  return true;
}

int gbit(int n, int wire) { return n - 1 - wire; }

// Declared condition number of a non-unitary op; unknown (<= 0, NaN) or
// singular counts as infinite.
double op_cond(const jq_op& op) {
  if (!(op.structure & JQ_S_NONUNITARY)) return 1.0;
  const double c = op.cond;
  return c >= 1.0 ? c : std::numeric_limits<double>::infinity();
}

}  // namespace

void build_program(const jq_plan_desc& d, int T_req, int RB, int RBb, HostProgram& out, int slot_budget,
                   int max_blocks) {
  const int n = d.n_wires;
  if (n < 1 || n > 40) throw PlanError(JQ_ERR_ARG, "n_wires must be in [1, 40]");
  // ---- fusion (simulation.py:214-240) ---------------------------------
  std::vector<HB> blocks;
  std::vector<int> last(n, -1);
  for (int oi = 0; oi < d.n_ops; ++oi) {
    const jq_op& op = d.ops[oi];
    if (op.nwires < 1 || op.nwires > JQ_MAX_WIRES)
      throw PlanError(JQ_ERR_UNSUPPORTED, "gates on more than 6 wires are not supported");
    std::vector<int> w(op.wires, op.wires + op.nwires);
    for (int x : w)
      if (x < 0 || x >= n) throw PlanError(JQ_ERR_ARG, "wire index out of range");
    std::vector<int> perm;
    if (op.structure & JQ_S_PERMUTATION) perm.assign(op.perm, op.perm + (1 << op.nwires));
    int i = last[w[0]];
    bool merge = i >= 0 && blocks[i].wires == w;
    if (merge)
      for (int x : w)
        if (last[x] != i) merge = false;
    if (merge) {
      HB& b = blocks[i];
      std::vector<int> np;
      if (!b.perm.empty() && !perm.empty()) {
        np.resize(perm.size());
        for (size_t r = 0; r < perm.size(); ++r) np[r] = b.perm[perm[r]];
      }
      b.perm = np;
      int s = 0;
      if ((b.structure & JQ_S_DIAGONAL) && (op.structure & JQ_S_DIAGONAL)) s |= JQ_S_DIAGONAL;
      if (!np.empty()) s |= JQ_S_PERMUTATION;
      if ((b.structure | op.structure) & JQ_S_NONUNITARY) s |= JQ_S_NONUNITARY;
      b.structure = s;
      b.cond *= op_cond(op);
      b.controls = std::min(b.controls, op.controls);
      for (int m = op.member_begin; m < op.member_end; ++m) b.members.push_back(m);
      continue;
    }
    HB b;
    b.wires = w;
    b.structure = op.structure;
    b.controls = op.controls;
    b.perm = perm;
    b.cond = op_cond(op);
    for (int m = op.member_begin; m < op.member_end; ++m) b.members.push_back(m);
    blocks.push_back(b);
    for (int x : w) last[x] = (int)blocks.size() - 1;
  }

  // ---- classify ---------------------------------------------------------
  const int maxk = std::min(RB, RBb);
  int pool = 0, bars = 0;
  for (HB& b : blocks) {
    b.k = (int)b.wires.size();
    const int full = 1 << b.k;
    b.nonunitary = (b.structure & JQ_S_NONUNITARY) != 0;
    for (int m : b.members) {
      const jq_member& mm = d.members[m];
      if (mm.param >= 0 || (mm.kind == JQ_M_MATRIX && (mm.flags & JQ_MF_GRAD))) b.grad = true;
      if (mm.param >= d.n_params) throw PlanError(JQ_ERR_ARG, "member parameter index out of range");
    }
    bool dense = false;
    // A non-unitary block is un-applied by its inverse in the backward sweep
    // only when it is well conditioned; otherwise it runs in a pass of its
    // own, whose backward reads the checkpointed input instead.
    const bool illcond = b.nonunitary && !(b.cond <= JQ_COND_MAX);
    if (b.structure & JQ_S_DIAGONAL) {
      b.kind = AK_DIAG;
      b.D = full;
      b.diagf = 1;
      if (illcond || b.k > 5 || (b.grad && b.k > 2)) b.wide = true;
    } else if (b.structure & JQ_S_PERMUTATION) {
      if (is_identity(b.perm)) {
        b.kind = AK_NOP;
      } else if (is_mcx(b.perm)) {
        b.kind = AK_MCX;
        b.targets = {b.wires[b.k - 1]};
        b.ctrls.assign(b.wires.begin(), b.wires.end() - 1);
      } else if (is_mcswap(b.perm)) {
        b.kind = AK_MCSWAP;
        b.targets = {b.wires[b.k - 2], b.wires[b.k - 1]};
        b.ctrls.assign(b.wires.begin(), b.wires.end() - 2);
      } else {
        dense = true;
        b.controls = 0;
      }
    } else {
      dense = true;
    }
    if (dense) {
      int c = std::max(0, std::min(b.controls, b.k - 1));
      int t = b.k - c;
      if (t == 1 && !illcond) {
        b.kind = AK_MC1;
        b.D = 2;
        // Products of Pauli rotations of the target (cos I - i sin P, P in
        // {X, Y, Z}, conjugated or not) and their angle derivatives all have
        // the form [[a, -b*], [b, a*]], so Re sum bar_ij dU_ij/dt only needs
        // bar00 + conj(bar11) and bar10 - conj(bar01).
        b.quat = !b.nonunitary;
        for (int m : b.members) {
          const jq_member& mm = d.members[m];
          b.quat = b.quat && mm.kind == JQ_M_ROT && !(mm.flags & JQ_MF_SUPEROP) && mm.word >= 1 && mm.word <= 3 &&
                   mm.nctrl == b.k - 1;
        }
      } else if (t == 2 && !illcond) {
        b.kind = AK_MC2;
        b.D = 4;
      } else if (c == 0 && b.k <= 3 && !b.grad && !illcond) {
        b.kind = AK_DENSEK;
        b.D = full;
      } else {
        b.wide = true;
      }
      if (!b.wide) {
        b.offs = full - b.D;
        b.targets.assign(b.wires.begin() + c, b.wires.end());
        b.ctrls.assign(b.wires.begin(), b.wires.begin() + c);
      }
    }
    if (b.wide) {
      // the full 2^k matrix (controls folded in), or its diagonal
      if (!b.diagf && b.k > WIDE_MAX_DENSE)
        throw PlanError(JQ_ERR_UNSUPPORTED, "dense blocks on more than 5 wires are not supported");
      b.kind = AK_WIDE;
      b.D = full;
      b.offs = 0;
      b.targets.clear();
      b.ctrls.clear();
    }
    if ((int)b.targets.size() > maxk) throw PlanError(JQ_ERR_UNSUPPORTED, "gate needs more register bits than the tile offers");
    if (b.kind == AK_MCX || b.kind == AK_MCSWAP || b.kind == AK_NOP) {
      b.grad = false;  // constant permutations carry no cotangent (simulation.py:626-629)
      if (b.nonunitary) throw PlanError(JQ_ERR_ARG, "permutation block marked non-unitary");
      continue;
    }
    const int elems = b.diagf ? b.D : b.D * b.D;
    b.mat = pool;
    pool += elems;
    if (b.nonunitary && !b.wide) {
      b.inv = pool;
      pool += elems;
    }
    if (b.grad) {
      b.bar = bars;
      b.bar_size = elems;
      bars += elems;
    }
  }

  // ---- tile bits ----------------------------------------------------------
  int T = T_req;
  if (T < RB) throw PlanError(JQ_ERR_ARG, "tile smaller than the register tile");
  const int ntiles_log = std::max(0, n - T);
  // Every tile holds the lowest global bits, so that loads and stores move
  // contiguous runs: 5 bits for complex64 (256-byte runs, 32 amplitudes) and
  // 6 for complex128 (1 KiB).  Fewer fixed bits leave more tile bits to the
  // gates and cut passes (the 24-qubit HEA: 24 passes with 7 fixed bits, 21
  // with 5); measured on MI355X, 5 is the best trade: 4 and 3 cost more in
  // access efficiency than they save in passes.
  const int lowdef = T >= 13 ? 5 : 6;
  if (d.low_bits != 0 && (d.low_bits < 3 || d.low_bits > 6)) throw PlanError(JQ_ERR_ARG, "low_bits must be 0 or 3..6");
  const int lowfix = std::min(n, std::min(T, d.low_bits ? d.low_bits : lowdef));
  // Global bits >= n are phantom (n < T): their amplitudes do not exist.

  out = HostProgram();
  out.T = T;
  out.RB = RB;
  out.n_wires = n;
  out.ntiles_log = ntiles_log;
  out.pool = pool;
  out.bars = bars;

  // block table and member table for the prep / chain kernels
  int bi_dev = 0;
  std::vector<int> dev_index(blocks.size(), -1);
  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    HB& b = blocks[bi];
    if (b.kind == AK_NOP || b.kind == AK_MCX || b.kind == AK_MCSWAP) continue;
    dev_index[bi] = bi_dev++;
    int rec[BW] = {0};
    rec[B_KIND] = b.kind;
    rec[B_K] = b.k;
    rec[B_D] = b.D;
    rec[B_MBEG] = (int)out.member_list.size();
    for (int m : b.members) out.member_list.push_back(m);
    rec[B_MEND] = (int)out.member_list.size();
    rec[B_MAT] = b.mat;
    rec[B_INV] = b.inv;
    rec[B_BAR] = b.bar;
    rec[B_DIAGF] = b.diagf;
    rec[B_OFFS] = b.offs;
    rec[B_QUAT] = b.kind == AK_MC1 && b.quat ? 1 : 0;
    out.block_table.insert(out.block_table.end(), rec, rec + BW);
    if (!b.wide && b.D > 8 && !b.diagf) throw PlanError(JQ_ERR_ARG, "internal: dense tile block wider than 3 wires");
    if (!b.wide && b.D > 32) throw PlanError(JQ_ERR_ARG, "internal: diagonal tile block wider than 5 wires");
  }
  out.n_dev_blocks = bi_dev;
  for (int m = 0; m < d.n_members; ++m) {
    const jq_member& mm = d.members[m];
    int rec[MW] = {mm.kind, mm.flags, mm.nctrl, mm.word, mm.param, mm.matrix, 0, 0};
    out.member_table.insert(out.member_table.end(), rec, rec + MW);
    out.member_values.push_back(mm.value);
  }
  // member generator wire count (k, or k/2 for a superoperator factor)
  for (HB& b : blocks)
    for (int m : b.members) {
      int kb = (d.members[m].flags & JQ_MF_SUPEROP) ? b.k / 2 : b.k;
      if ((d.members[m].flags & JQ_MF_SUPEROP) && (b.k % 2)) throw PlanError(JQ_ERR_ARG, "superoperator on an odd wire count");
      out.member_table[m * MW + M_KB] = kb;
    }

  // ---- passes ---------------------------------------------------------------
  std::vector<int> remaining;
  for (size_t bi = 0; bi < blocks.size(); ++bi)
    if (blocks[bi].kind != AK_NOP) remaining.push_back((int)bi);

  struct Pass {
    std::vector<int> tile;  // global bits ascending
    std::vector<int> blocks;
    bool wide = false;      // one AK_WIDE block
  };
  std::vector<Pass> passes;
  auto low_bits = [&]() {
    std::vector<int> v;
    for (int g = 0; g < lowfix; ++g) v.push_back(g);
    return v;
  };
  while (!remaining.empty()) {
    std::vector<char> in_tile(n + T + 1, 0), blocked(n, 0);
    int tcount = 0;
    if (n <= T) {
      for (int g = 0; g < n; ++g) in_tile[g] = 1;
      tcount = n;
    } else {
      for (int g : low_bits()) in_tile[g] = 1, ++tcount;
    }
    Pass P;
    std::vector<int> left;
    int pslots = 0;  // cotangent reals flushed by this pass
    if (blocks[remaining[0]].wide) {  // a wide block runs alone
      P.wide = true;
      P.blocks.push_back(remaining[0]);
      left.assign(remaining.begin() + 1, remaining.end());
      for (int g = 0; g < T; ++g) P.tile.push_back(g);
      passes.push_back(P);
      remaining = left;
      continue;
    }
    for (int bi : remaining) {
      HB& b = blocks[bi];
      if (max_blocks > 0 && (int)P.blocks.size() >= max_blocks) {  // the pass is full: the rest waits, in order
        left.push_back(bi);
        continue;
      }
      bool blk = b.wide;  // deferred to a pass of its own
      if (blk) {
        for (int w : b.wires) blocked[gbit(n, w)] = 1;
        left.push_back(bi);
        continue;
      }
      for (int w : b.wires)
        if (blocked[gbit(n, w)]) blk = true;
      const int reals = b.grad ? b.bar_size * 2 : 0;
      if (!blk && pslots + reals <= slot_budget) {
        int need = 0;
        for (int w : b.targets)
          if (!in_tile[gbit(n, w)]) ++need;
        if (tcount + need <= T) {
          for (int w : b.targets)
            if (!in_tile[gbit(n, w)]) in_tile[gbit(n, w)] = 1, ++tcount;
          P.blocks.push_back(bi);
          pslots += reals;
          continue;
        }
      }
      for (int w : b.wires) blocked[gbit(n, w)] = 1;
      left.push_back(bi);
    }
    if (P.blocks.empty()) throw PlanError(JQ_ERR_UNSUPPORTED, "scheduler made no progress");
    // pad the tile to T bits: real bits first (lowest unused), then phantom bits
    for (int g = 0; g < n && tcount < T; ++g)
      if (!in_tile[g]) in_tile[g] = 1, ++tcount;
    for (int g = n; tcount < T; ++g) in_tile[g] = 1, ++tcount;
    for (int g = 0; g < (int)in_tile.size(); ++g)
      if (in_tile[g]) P.tile.push_back(g);
    passes.push_back(P);
    remaining = left;
  }
  if (passes.empty() || passes.back().wide) {
    // empty circuit: one pass with no gates (copies psi0); after a final
    // wide block, a gate-free tile pass that measures and seeds the adjoint
    Pass P;
    for (int g = 0; g < T; ++g) P.tile.push_back(g);
    passes.push_back(P);
  }

  // ---- cotangent slots (shared by the forward and backward programs) -------
  const int ntiles = 1 << ntiles_log;
  std::vector<int> slot_of(blocks.size(), -1);
  std::vector<int> pass_nslots(passes.size(), 0), pass_slot_base(passes.size(), 0);
  {
    int slot_base = 0;
    for (size_t p = 0; p < passes.size(); ++p) {
      int nslots = 0;
      for (int bi : passes[p].blocks) {
        HB& b = blocks[bi];
        if (!b.grad) continue;
        slot_of[bi] = nslots;
        out.slot_map.push_back({(int)p, nslots, b.bar * 2, b.bar_size * 2, slot_base});
        nslots += b.bar_size * 2;
      }
      pass_nslots[p] = nslots;
      pass_slot_base[p] = slot_base;
      slot_base += nslots * ntiles;
    }
    out.partial_reals = slot_base;
  }

  // ---- emit (once per register-tile size) -----------------------------------
  auto emit = [&](const int RB, std::vector<int32_t>& prog, int& total_sub) {
  prog.assign(HDR_WORDS + passes.size(), 0);
  prog[0] = PROG_MAGIC;
  prog[1] = (int)passes.size();
  prog[2] = T;
  prog[3] = RB;
  prog[4] = n;
  prog[5] = bi_dev;
  prog[6] = bars * 2;
  prog[8] = ntiles;
  total_sub = 0;
  for (size_t p = 0; p < passes.size(); ++p) {
    Pass& P = passes[p];
    prog[HDR_WORDS + p] = (int)prog.size();
    const size_t rec = prog.size();
    prog.resize(rec + 4 + T, 0);
    int flags = 0;
    if (p == 0) flags |= PF_FIRST;
    if (p + 1 == passes.size()) flags |= PF_LAST;
    if (P.wide) flags |= PF_WIDE;
    prog[rec + 1] = flags;
    std::vector<int> tloc(n + T + 1, -1);  // global bit -> tile-local bit
    for (int j = 0; j < T; ++j) {
      prog[rec + 4 + j] = P.tile[j];
      tloc[P.tile[j]] = j;
    }
    // substeps
    std::vector<std::vector<int32_t>> subwords;
    std::vector<int> rem = P.wide ? std::vector<int>() : P.blocks;
    int nsub = 0;
    if (P.wide) {  // one substep (default layout) holding the wide block's record
      HB& b = blocks[P.blocks[0]];
      std::vector<int32_t> sw(1 + T, 0);
      sw[0] = 1;
      for (int i = 0; i < RB; ++i) sw[1 + i] = T - 1 - i;
      for (int j = 0; j < T - RB; ++j) sw[1 + RB + j] = j;
      int g[GW];
      for (int& x : g) x = 0;
      g[G_KIND] = AK_WIDE;
      g[G_BLOCK] = dev_index[P.blocks[0]];
      g[G_RA] = b.diagf;
      g[G_MAT] = b.mat;
      g[G_INV] = -1;
      g[G_SLOT] = slot_of[P.blocks[0]];
      g[G_K] = b.k;
      for (int j = 0; j < b.k; ++j) g[G_DW0 + j / 4] |= gbit(n, b.wires[j]) << (8 * (j % 4));
      sw.insert(sw.end(), g, g + GW);
      subwords.push_back(sw);
      nsub = 1;
    }
    std::vector<int> prev_reg;
    while (!rem.empty()) {
      std::vector<int> reg;  // tile-local bits in registers, by register position
      std::vector<char> blocked(n, 0);
      std::vector<int> sub, left;
      int forced_k = 0;
      for (int bi : rem) {
        HB& b = blocks[bi];
        bool blk = false;
        for (int w : b.wires)
          if (blocked[gbit(n, w)]) blk = true;
        bool ok = false;
        if (!blk) {
          if (b.kind == AK_DIAG) {
            ok = true;
          } else if (b.kind == AK_DENSEK) {
            if (reg.empty() && forced_k == 0) {
              for (int j = b.k - 1; j >= 0; --j) reg.push_back(tloc[gbit(n, b.wires[j])]);
              forced_k = b.k;
              ok = true;
            }
          } else {
            int need = 0;
            for (int w : b.targets)
              if (std::find(reg.begin(), reg.end(), tloc[gbit(n, w)]) == reg.end()) ++need;
            if ((int)reg.size() + need <= RB) {
              for (int w : b.targets) {
                int tb = tloc[gbit(n, w)];
                if (std::find(reg.begin(), reg.end(), tb) == reg.end()) reg.push_back(tb);
              }
              ok = true;
            }
          }
        }
        if (ok) {
          sub.push_back(bi);
        } else {
          for (int w : b.wires) blocked[gbit(n, w)] = 1;
          left.push_back(bi);
        }
      }
      // Register layout of this substep.  Passes load and store under the
      // canonical layout (register bit 0 <-> tile bit 0, the other register
      // bits <-> the top tile bits, lanes <-> the low tile bits); a substep
      // keeps the previous layout, else the canonical one, whenever its
      // targets fit, so the LDS re-layout between substeps is skipped.
      std::vector<int> canon(RB);
      canon[0] = 0;
      for (int b2 = 1; b2 < RB; ++b2) canon[b2] = T - RB + b2;
      auto covers = [&](const std::vector<int>& lay) {
        for (int tb : reg)
          if (std::find(lay.begin(), lay.end(), tb) == lay.end()) return false;
        return true;
      };
      std::vector<char> used(T, 0);
      if (forced_k == 0) {
        std::vector<int> lay;
        if (prev_reg.empty()) prev_reg = canon;
        if (covers(prev_reg)) {
          lay = prev_reg;
        } else if (covers(canon)) {
          lay = canon;
        } else {
          lay = canon;
          std::vector<char> keep(RB, 0);
          for (int b2 = 0; b2 < RB; ++b2)
            if (std::find(reg.begin(), reg.end(), canon[b2]) != reg.end()) keep[b2] = 1;
          int pos = RB - 1;
          for (int tb : reg) {
            if (std::find(canon.begin(), canon.end(), tb) != canon.end()) continue;
            while (pos >= 0 && keep[pos]) --pos;
            lay[pos] = tb;
            keep[pos] = 1;
          }
        }
        reg = lay;
      } else {
        for (int tb : reg) used[tb] = 1;
        for (int tb : canon)
          if ((int)reg.size() < RB && !used[tb]) reg.push_back(tb), used[tb] = 1;
        for (int tb = T - 1; tb >= 0 && (int)reg.size() < RB; --tb)
          if (!used[tb]) reg.push_back(tb), used[tb] = 1;
      }
      std::fill(used.begin(), used.end(), 0);
      for (int tb : reg) used[tb] = 1;
      prev_reg = reg;
      std::vector<int32_t> sw(1 + T, 0);
      sw[0] = (int)sub.size();
      std::vector<int> regpos(T, -1), thrpos(T, -1);
      for (int i = 0; i < RB; ++i) {
        sw[1 + i] = reg[i];
        regpos[reg[i]] = i;
      }
      int tj = 0;
      for (int tb = 0; tb < T; ++tb)
        if (!used[tb]) {
          sw[1 + RB + tj] = tb;
          thrpos[tb] = tj;
          ++tj;
        }
      for (int bi : sub) {
        HB& b = blocks[bi];
        int g[GW];
        for (int& x : g) x = 0;
        g[G_KIND] = b.kind;
        g[G_BLOCK] = dev_index[bi];
        g[G_RA] = b.targets.size() > 0 ? regpos[tloc[gbit(n, b.targets[0])]] : 0;
        g[G_RB] = b.targets.size() > 1 ? regpos[tloc[gbit(n, b.targets[1])]] : 0;
        if (b.kind == AK_DENSEK) g[G_RA] = b.k;
        uint32_t creg = 0, cthr = 0;
        uint64_t cout = 0;
        for (int w : b.ctrls) {
          int gb = gbit(n, w);
          int tb = tloc[gb];
          if (tb >= 0 && regpos[tb] >= 0)
            creg |= 1u << regpos[tb];
          else if (tb >= 0)
            cthr |= 1u << tb;
          else
            cout |= 1ull << gb;
        }
        g[G_CREG] = (int)creg;
        g[G_CTHR] = (int)cthr;
        g[G_COUT_LO] = (int)(uint32_t)(cout & 0xffffffffu);
        g[G_COUT_HI] = (int)(uint32_t)(cout >> 32);
        g[G_MAT] = b.mat;
        g[G_INV] = b.inv;
        g[G_SLOT] = slot_of[bi];
        g[G_K] = b.k;
        g[G_QUAT] = b.kind == AK_MC1 && b.quat ? 1 : 0;
        if (b.kind == AK_DIAG) {
          for (int j = 0; j < b.k; ++j) {
            int gb = gbit(n, b.wires[j]);
            int tb = tloc[gb];
            int loc;
            if (tb >= 0 && regpos[tb] >= 0)
              loc = (DL_REG << 6) | regpos[tb];
            else if (tb >= 0)
              loc = (DL_THR << 6) | tb;
            else
              loc = (DL_OUT << 6) | gb;
            g[G_DW0 + j / 4] |= loc << (8 * (j % 4));
          }
        }
        sw.insert(sw.end(), g, g + GW);
      }
      subwords.push_back(sw);
      ++nsub;
      rem = left;
    }
    if (nsub == 0) {  // gate-free pass: one substep with the default layout
      std::vector<int32_t> sw(1 + T, 0);
      for (int i = 0; i < RB; ++i) sw[1 + i] = T - 1 - i;
      for (int j = 0; j < T - RB; ++j) sw[1 + RB + j] = j;
      subwords.push_back(sw);
      nsub = 1;
    }
    const size_t offs = prog.size();
    prog.resize(offs + nsub, 0);
    for (int k = 0; k < nsub; ++k) {
      prog[offs + k] = (int)prog.size();
      prog.insert(prog.end(), subwords[k].begin(), subwords[k].end());
    }
    prog[rec] = nsub;
    prog[rec + 2] = pass_nslots[p];
    prog[rec + 3] = pass_slot_base[p];
    total_sub += nsub;
  }
  prog[7] = out.partial_reals;
  };
  int sub_f = 0, sub_b = 0;
  emit(RB, out.words, sub_f);
  emit(RBb, out.words_b, sub_b);
  out.n_passes = (int)passes.size();
  out.n_substeps = sub_f;
  out.ntiles = ntiles;
}

// Local affine map of a permutation op: basis state |c> goes to |pi(c)>
// with pi = perm^-1 (perm maps output rows to input columns).  Returns false
// unless pi(c) = A c ^ b for all c; cols[j] = A e_j.
static bool affine_perm(const jq_op& op, std::vector<uint32_t>& cols, uint32_t& b) {
  const int k = op.nwires, full = 1 << k;
  std::vector<int> inv(full, -1);
  for (int r = 0; r < full; ++r) {
    const int c = op.perm[r];
    if (c < 0 || c >= full || inv[c] >= 0) return false;
    inv[c] = r;
  }
  b = (uint32_t)inv[0];
  cols.assign(k, 0);
  for (int j = 0; j < k; ++j) cols[j] = (uint32_t)inv[1 << j] ^ b;
  for (int c = 0; c < full; ++c) {
    uint32_t y = b;
    for (int j = 0; j < k; ++j)
      if ((c >> j) & 1) y ^= cols[j];
    if (y != (uint32_t)inv[c]) return false;
  }
  return true;
}

void fold_tail(const jq_plan_desc& d, std::vector<jq_op>& kept, TailFold& f) {
  const int n = d.n_wires;
  f = TailFold();
  f.rows.resize(n);
  for (int w = 0; w < n; ++w) f.rows[w] = 1ull << w;
  kept.assign(d.ops, d.ops + d.n_ops);
  if (n > 64) return;
  auto wire_mask = [&](const jq_op& op) {
    uint64_t m = 0;
    for (int i = 0; i < op.nwires; ++i)
      if (op.wires[i] >= 0 && op.wires[i] < n) m |= 1ull << op.wires[i];
    return m;
  };
  auto foldable = [&](const jq_op& op) {
    if (op.nwires < 1 || op.nwires > JQ_MAX_WIRES || (op.structure & JQ_S_NONUNITARY)) return false;
    for (int i = 0; i < op.nwires; ++i)
      if (op.wires[i] < 0 || op.wires[i] >= n) return false;
    const bool perm = (op.structure & JQ_S_PERMUTATION) != 0;
    const bool diag = (op.structure & JQ_S_DIAGONAL) != 0;
    if (!perm && !diag) return false;
    for (int m = op.member_begin; m < op.member_end; ++m) {
      if (m < 0 || m >= d.n_members) return false;
      const jq_member& mm = d.members[m];
      if (mm.param >= 0 || (mm.flags & (JQ_MF_GRAD | JQ_MF_SUPEROP))) return false;
      // a diagonal op is unitary when it is built from rotations, phases and
      // matrices the recording vouches for
      if (!perm && mm.kind == JQ_M_MATRIX && !(mm.flags & JQ_MF_UNITARY)) return false;
    }
    if (perm) {
      std::vector<uint32_t> cols;
      uint32_t b;
      return affine_perm(op, cols, b);
    }
    return true;
  };
  // walk back from the end: an op folds when no kept op comes after it on
  // its wires (ops on disjoint wires commute)
  std::vector<char> folded(d.n_ops, 0);
  uint64_t blocked = 0;
  int nf = 0;
  for (int i = d.n_ops - 1; i >= 0; --i) {
    const jq_op& op = d.ops[i];
    const uint64_t wm = wire_mask(op);
    if (!(wm & blocked) && foldable(op)) {
      folded[i] = 1;
      ++nf;
    } else {
      blocked |= wm;
    }
  }
  if (nf == 0 || nf == d.n_ops) return;
  // compose the folded permutations in circuit order
  for (int i = 0; i < d.n_ops; ++i) {
    if (!folded[i] || !(d.ops[i].structure & JQ_S_PERMUTATION)) continue;
    const jq_op& op = d.ops[i];
    const int k = op.nwires;
    std::vector<uint32_t> cols;
    uint32_t b;
    affine_perm(op, cols, b);
    // local bit j is wire op.wires[k - 1 - j] (the first wire is the most
    // significant bit of the op's index)
    std::vector<uint64_t> nr(k);
    std::vector<int> nflip(k);
    for (int li = 0; li < k; ++li) {
      uint64_t r = 0;
      int fl = (b >> li) & 1;
      for (int j = 0; j < k; ++j)
        if ((cols[j] >> li) & 1) {
          const int w = op.wires[k - 1 - j];
          r ^= f.rows[w];
          fl ^= (int)((f.flip >> w) & 1);
        }
      nr[li] = r;
      nflip[li] = fl;
    }
    for (int li = 0; li < k; ++li) {
      const int w = op.wires[k - 1 - li];
      f.rows[w] = nr[li];
      f.flip = (f.flip & ~(1ull << w)) | ((uint64_t)nflip[li] << w);
    }
  }
  kept.clear();
  for (int i = 0; i < d.n_ops; ++i)
    if (!folded[i]) kept.push_back(d.ops[i]);
  f.n_folded = nf;
}

}  // namespace jq

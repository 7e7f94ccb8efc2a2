// Device program format shared by the planner (plan.cpp) and the kernels.
//
// A plan is one circuit structure.  Its lowered ops are fused into blocks
// (same-wire runs, jaqsi/simulation.py:214-240) and the blocks are scheduled
// into *passes*: a pass streams the vector once through HBM in tiles of 2^T
// amplitudes; inside a tile the amplitudes live in registers, 2^RB per
// thread, and a *substep* is a register layout of the tile under which a run
// of blocks is applied with no data movement.  Between substeps the tile is
// re-laid out through LDS.
//
// Program words (int32):
//   header  [0] magic  [1] n_passes  [2] T  [3] RB  [4] n_wires  [5] n_blocks
//           [6] grad reals per circuit (bars)  [7] partial reals per circuit
//           [8] expval partial tiles (ntiles)  [9..15] reserved
//           [16 .. 16+n_passes) offset of each pass record
//   pass    [0] n_sub [1] flags [2] nslots (grad reals per tile) [3] slot_base
//           [4 .. 4+T) tile_bits: global bit of tile bit j (ascending)
//           then n_sub substep records
//   substep [0] n_gates  [1 .. 1+RB) reg_src (tile bit of register bit i)
//           [1+RB .. 1+T) thr_src (tile bit of thread bit j)
//           then n_gates gate records of GW words
#pragma once
#ifndef __HIPCC_RTC__
#include <stdint.h>
#endif

namespace jq {

constexpr int32_t PROG_MAGIC = 0x4a514131;  // "JQA1"
constexpr int HDR_WORDS = 16;

// pass flags
constexpr int PF_FIRST = 1;  // first pass of the circuit (reads psi0 or |0>)
constexpr int PF_LAST = 2;   // last pass (computes expvals; seeds lambda in the backward)
constexpr int PF_WIDE = 4;   // a pass of one wide block (AK_WIDE), run by the wide-block kernels

// apply kinds
constexpr int AK_NOP = 0;
constexpr int AK_MC1 = 1;     // 2x2 on target reg bit ra, controlled by masks
constexpr int AK_MC2 = 2;     // 4x4 on reg bits (ra = first target, rb = second)
constexpr int AK_MCX = 3;     // X on reg bit ra under controls
constexpr int AK_MCSWAP = 4;  // swap reg bits ra, rb under controls
constexpr int AK_DIAG = 5;    // diagonal over k wires located anywhere
constexpr int AK_DENSEK = 6;  // dense 2^k on reg bits k-1..0 (wire j <-> reg bit k-1-j)
// A block applied in a pass of its own, by the wide-block kernels: dense on
// up to 5 wires (controls folded into the 2^k matrix) or diagonal on up to
// 6, with its cotangent reduced over the whole vector; its backward reads
// the checkpointed input state, so it is never inverted.  Record words:
// G_BLOCK, G_MAT, G_SLOT, G_K, G_RA = 1 for diagonal storage, and
// G_DW0 / G_DW0+1 = global bit of wire j in byte j (wire 0 = most
// significant bit of the matrix index).
constexpr int AK_WIDE = 7;
constexpr int WIDE_MAX_DENSE = 5;
constexpr int WIDE_MAX_DIAG = 6;

// gate record words
constexpr int GW = 16;
constexpr int G_KIND = 0;
constexpr int G_BLOCK = 1;
constexpr int G_RA = 2;
constexpr int G_RB = 3;
constexpr int G_CREG = 4;     // control mask over register-index bits
constexpr int G_CTHR = 5;     // control mask over tile-local bits held by thread bits
constexpr int G_COUT_LO = 6;  // control mask over global bits outside the tile
constexpr int G_COUT_HI = 7;
constexpr int G_MAT = 8;      // element offset of the block matrix in the pool
constexpr int G_INV = 9;      // element offset of the inverse (non-unitary), -1 otherwise
constexpr int G_SLOT = 10;    // grad slot (reals, within the pass tile area), -1 none
constexpr int G_K = 11;       // DIAG / DENSEK: number of wires
constexpr int G_DW0 = 12;     // DIAG wire locators (4 words, 8 bits each: [type:2][pos:6])
constexpr int G_QUAT = 15;    // MC1: 1 when every factor is a Pauli rotation of the target (quaternion form)
// DIAG locator types
constexpr int DL_REG = 0;     // pos = register bit
constexpr int DL_THR = 1;     // pos = tile-local bit (held by thread bits)
constexpr int DL_OUT = 2;     // pos = global bit outside the tile

// block table (per plan, int32 records of BW words) for prep / chain kernels
constexpr int BW = 12;
constexpr int B_KIND = 0;
constexpr int B_K = 1;        // wires of the block
constexpr int B_D = 2;        // effective matrix dimension (active block) or diag length
constexpr int B_MBEG = 3;     // member range
constexpr int B_MEND = 4;
constexpr int B_MAT = 5;      // pool offset of the effective matrix
constexpr int B_INV = 6;      // pool offset of its inverse, -1
constexpr int B_BAR = 7;      // offset (complex elements) of the cotangent in bars, -1
constexpr int B_DIAGF = 8;    // 1 if diagonal storage
constexpr int B_OFFS = 9;     // active-block offset inside the full 2^k matrix
constexpr int B_QUAT = 10;    // 1: MC1 block of Pauli rotations of its target (G_QUAT)

// member table (int32 records of MW words + the double angle)
constexpr int MW = 8;
constexpr int M_KIND = 0;
constexpr int M_FLAGS = 1;
constexpr int M_NCTRL = 2;
constexpr int M_WORD = 3;
constexpr int M_PARAM = 4;
constexpr int M_MATRIX = 5;
constexpr int M_KB = 6;       // wires of the factor's generator (k, or k/2 for a superoperator)

}  // namespace jq

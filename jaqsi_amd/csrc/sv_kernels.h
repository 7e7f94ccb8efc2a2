// Launch interface of the tile-pass kernels (sv_kernels.hip).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace jq {

// One plan's circuits inside a multi-plan launch (circuits of several
// structures -- e.g. uneven depths -- share every pass launch).  Pass q of a
// group reads its input from the previous pass output and writes its own:
// outputs of passes 0 .. n_passes-2 live in `ckpt` ([pass][circuit][amp]),
// the last one in `state` ([circuit][amp]).
struct SvGroup {
  const int* prog_f;     // forward program (device)
  const int* prog_b;     // backward program (device)
  const void* mats;      // block matrices [batch][pool_stride]
  int64_t pool_stride;
  const void* psi0;      // initial states or nullptr
  int64_t psi0_stride;
  void* ckpt;
  void* state;
  void* lam;             // [circuit][amp]
  double* ev_part;
  double* grad_part;
  int64_t part_stride;
  const double* ct;      // [circuit][n_obs]
  int n_passes;
  int circ_begin;        // first grid row of the group
  int batch;
  int pad;
};

struct SvArgs {
  const int* prog;     // scheduled program (device)
  int pass;            // pass index of this launch
  int n_wires;
  uint64_t vec_len;    // 2^n_wires
  const void* mats;    // block matrices [batch][pool_stride] (plan dtype)
  int64_t pool_stride;
  int rows;            // persistent backward: circuits of the launch (grid is flat)
  int64_t pool_off;    // element offset added to every matrix reference of the launch
                       // (0 for the program interpreter; plan-specialised kernels are
                       // shared by passes that differ only by this offset)
  const void* src;     // input vectors; nullptr => |0...0>
  int64_t src_stride;  // elements between circuits (0 = broadcast)
  void* dst;           // output vectors (forward) / running psi (backward, may be nullptr)
  int64_t dst_stride;
  void* lam;           // backward: lambda vectors in/out
  int64_t lam_stride;
  int obs_kind;
  int n_obs;
  const void* obs;       // device ObsTable (observable wires and diagonals)
  const double* ct;      // [batch][n_obs] cotangents (device)
  double* ev_part;       // [batch][ntiles][n_obs]
  double* grad_part;     // [batch][part_stride]
  int64_t part_stride;
  // multi-plan launches: when `groups` is set, grid row y belongs to the
  // group whose circ_begin is the largest <= y; the launch runs step `step`
  // of direction `dir` (0 forward: pass = step; 1 backward: pass =
  // n_passes - 1 - step) and every per-circuit pointer above comes from the
  // group record.
  const SvGroup* groups;
  int n_groups;
  int step;
  int dir;
  // backward: the caller has put the cotangent of the final vector in
  // `lam`; the last pass reads it instead of seeding it from observables
  int lam_given;
};

// Observable table: global bit and diagonal (d0, d1) of each observable wire.
struct ObsTable {
  int gbit[64];         // single-wire kinds: global bit of the observable wire
  double d[128];        // (d0, d1) per observable
  uint64_t mask[64];    // JQ_OBS_SV_PARITY: global-bit mask of the Pauli-Z word
};

#ifndef __HIPCC_RTC__  // host launch interface (absent from run-time compiled kernels)
// Copy a host-built table to device memory on the stream (a tiny kernel, so
// the call stays asynchronous and capturable).
hipError_t launch_obs_upload(const ObsTable& t, void* dst, hipStream_t st);

// Copy n host bytes to device memory on the stream through kernel arguments
// (asynchronous and capturable; for small tables).
hipError_t launch_bytes_upload(const void* src, size_t n, void* dst, hipStream_t st);

// which: 0 forward, 1 backward, +2 full variant (4x4 / dense 3-wire blocks),
// +4 backward with the wide register tile (complex64 only)
hipError_t launch_sv_f32(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st);
hipError_t launch_sv_f64(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st);

inline hipError_t launch_sv(int which, int dtype, int T, const SvArgs& a, int ntiles, int batch, size_t smem,
                            hipStream_t st) {
  return dtype == 0 ? launch_sv_f32(which, T, a, ntiles, batch, smem, st)
                    : launch_sv_f64(which, T, a, ntiles, batch, smem, st);
}

#endif

}  // namespace jq

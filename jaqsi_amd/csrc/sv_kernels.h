// Launch interface of the tile-pass kernels (sv_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jq {

struct SvArgs {
  const int* prog;     // scheduled program (device)
  int pass;            // pass index of this launch
  int n_wires;
  uint64_t vec_len;    // 2^n_wires
  const void* mats;    // block matrices [batch][pool_stride] (plan dtype)
  int64_t pool_stride;
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
};

// Observable table: global bit and diagonal (d0, d1) of each observable wire.
struct ObsTable {
  int gbit[64];
  double d[128];
};

// Copy a host-built table to device memory on the stream (a tiny kernel, so
// the call stays asynchronous and capturable).
hipError_t launch_obs_upload(const ObsTable& t, void* dst, hipStream_t st);

// which: 0 forward pass, 1 backward pass, 2 fused single-tile forward+backward
hipError_t launch_sv_f32(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st);
hipError_t launch_sv_f64(int which, int T, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st);

inline hipError_t launch_sv(int which, int dtype, int T, const SvArgs& a, int ntiles, int batch, size_t smem,
                            hipStream_t st) {
  return dtype == 0 ? launch_sv_f32(which, T, a, ntiles, batch, smem, st)
                    : launch_sv_f64(which, T, a, ntiles, batch, smem, st);
}

}  // namespace jq

// Launch interface of the small kernels (prep.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jq {

struct PrepArgs {
  int n_blocks, batch;
  const int* block_table;
  const int* member_list;
  const int* member_table;
  const double* member_values;
  const double* angles;     // [batch][n_params]
  int n_params;
  const void* matrices;     // [batch][n_matrix_elems] complex128, may be null
  int64_t n_matrix_elems;
  void* pool;               // [batch][pool_stride] plan dtype
  int64_t pool_stride;
  int class_mask;           // block classes present (block_class), one kernel instance each
  const int* wg_blocks;     // blocks of the workgroup classes (block_class >= 3)
  int n_wg;
};

// Block classes of the prep / chain kernels: 0 diagonal <= 4, 1 diagonal
// <= 32, 2 dense 2x2 (one thread per block, matrices in registers);
// 3 dense 4x4, 4 dense 8x8, 5 dense 16x16 / 32x32 and diagonals of 64
// (one workgroup per block, one matrix entry per thread, matrices in LDS).
__host__ __device__ inline int block_class(bool diagf, int D) {
  if (diagf) return D <= 4 ? 0 : D <= 32 ? 1 : 5;
  return D == 2 ? 2 : D == 4 ? 3 : D == 8 ? 4 : 5;
}
constexpr int WG_FIRST_CLASS = 3;

struct ReduceArgs {
  const int* slot_map;      // n_slots x {base, nslots, local, bar, size}, sorted by bar
  int n_slots;
  const double* partials;   // [batch][part_stride]
  int64_t part_stride;
  int ntiles;
  double* bars;             // [batch][n_reals]
  int n_reals;
  int batch;
};

struct ChainArgs {
  int n_blocks, batch;
  const int* block_table;
  const int* member_list;
  const int* member_table;
  const double* member_values;
  const double* angles;
  int n_params;
  const void* matrices;
  int64_t n_matrix_elems;
  const double* bars;       // complex128 cotangents [batch][bars_stride]
  int64_t bars_stride;      // complex elements
  double* grad_angles;      // [batch][n_params], pre-zeroed
  void* grad_matrices;      // [batch][n_matrix_elems] complex128, pre-zeroed, may be null
  int class_mask;           // block classes present (block_class)
  int quat_t;               // 1: B_QUAT blocks hold (t_x, t_y, t_z) = Im Tr(bar^T U sigma_k) in
                            // (bar00.re, bar00.im, bar10.re) (the plan-specialised backward's form)
  const int* wg_blocks;     // blocks of the workgroup classes (block_class >= 3)
  int n_wg;
  int max_members2;         // most members of a 2x2 (class 2) block
};

struct ExpvalArgs {
  int n_qubits, k, density;
  const void* state;
  int64_t stride;  // elements between states
  int64_t wmask;   // global bits of the observable wires
  int gbits[5];    // global bit of wire j of the observable
  const void* mat; // complex128 2^k x 2^k (device)
  double* out;     // [batch], pre-zeroed
};

hipError_t launch_prep(int dtype, const PrepArgs& a, hipStream_t st);
hipError_t launch_reduce_grad(const ReduceArgs& a, hipStream_t st);
hipError_t launch_reduce_ev(const double* part, double* out, int ntiles, int n_obs, int batch, hipStream_t st);
// Zero `bytes` (a multiple of 4, 4-byte aligned) with a kernel.
hipError_t launch_zero(void* p, size_t bytes, hipStream_t st);
hipError_t launch_chain(const ChainArgs& a, hipStream_t st);
hipError_t launch_probs(int dtype, const void* s, void* p, int64_t total, hipStream_t st);
hipError_t launch_outer(int dtype, const void* s, void* rho, int64_t dim, int64_t batch, hipStream_t st);
hipError_t launch_dm_diag(int dtype, const void* rho, void* p, int64_t dim, int64_t batch, hipStream_t st);
hipError_t launch_expval_matrix(int dtype, const ExpvalArgs& a, int batch, hipStream_t st);

}  // namespace jq

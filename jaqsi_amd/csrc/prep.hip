// Small kernels around the tile passes:
//   prep    -- block matrices from the angle vector (the fused products of
//              _fuse, simulation.py:214-240, evaluated on the device in FP64)
//   reduce  -- per-tile cotangent / expectation partial sums -> FP64 totals
//   chain   -- block cotangents -> angle gradients and explicit-matrix
//              cotangents (the chain rule JAX applies through _rot_matrix,
//              simulation.py:555-558 / gateset.py:399-405)
//   probs / outer / dm_diag / expval_matrix -- measurement outputs
//              (simulation.py:735-857)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "jaqsi_amd.h"
#include "prep.h"
#include "sv_kernels.h"
#include "program.h"

namespace jq {

struct cd {
  double x, y;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd cconj(cd a) { return {a.x, -a.y}; }

// entry (a, b) of a single-qubit Pauli (code I0 X1 Y2 Z3)
__device__ __forceinline__ cd pauli_entry(int code, int a, int b) {
  switch (code) {
    case 0: return {a == b ? 1.0 : 0.0, 0.0};
    case 1: return {a != b ? 1.0 : 0.0, 0.0};
    case 2: return a == b ? cd{0.0, 0.0} : (a == 0 ? cd{0.0, -1.0} : cd{0.0, 1.0});
    default: return {a == b ? (a ? -1.0 : 1.0) : 0.0, 0.0};
  }
}

// generator entry g[r][c] (or its theta-derivative) of a member over kb wires
__device__ cd gen_entry(const int* mt, double theta, const cd* xm, int kb, int r, int c, bool deriv) {
  const int kind = mt[M_KIND];
  const int dim = 1 << kb;
  if (kind == JQ_M_MATRIX) {
    if (deriv) return {0.0, 0.0};
    return xm[mt[M_MATRIX] + r * dim + c];
  }
  const int nc = mt[M_NCTRL];
  const int t = kb - nc;
  const int s = dim - (1 << t);
  if (r < s || c < s) return {(!deriv && r == c) ? 1.0 : 0.0, 0.0};
  const int rr = r - s, cc = c - s;
  if (kind == JQ_M_PHASE) {
    if (rr == (1 << t) - 1 && cc == rr) {
      double sn, cs;
      sincos(theta, &sn, &cs);
      return deriv ? cd{-sn, cs} : cd{cs, sn};
    }
    return {(!deriv && rr == cc) ? 1.0 : 0.0, 0.0};
  }
  // rotation: cos(t/2) I - i sin(t/2) P
  const int word = mt[M_WORD];
  cd p = {1.0, 0.0};
  for (int j = 0; j < t; ++j) {
    const int code = (word >> (2 * (t - 1 - j))) & 3;
    p = cmul(p, pauli_entry(code, (rr >> (t - 1 - j)) & 1, (cc >> (t - 1 - j)) & 1));
  }
  double sn, cs;
  sincos(0.5 * theta, &sn, &cs);
  const double dlt = rr == cc ? 1.0 : 0.0;
  if (!deriv) return {cs * dlt + sn * p.y, -sn * p.x};  // cs*d - i*sn*p
  return {-0.5 * sn * dlt + 0.5 * cs * p.y, -0.5 * cs * p.x};
}

// entry (r, c) of the member's full 2^k matrix (k block wires), or its derivative
__device__ cd member_entry(const int* mt, double theta, const cd* xm, int r, int c, bool deriv) {
  const int flags = mt[M_FLAGS];
  const int kb = mt[M_KB];
  cd v;
  if (flags & JQ_MF_SUPEROP) {
    const int msk = (1 << kb) - 1;
    const int r1 = r >> kb, r2 = r & msk, c1 = c >> kb, c2 = c & msk;
    const cd a = gen_entry(mt, theta, xm, kb, r1, c1, false);
    const cd b = cconj(gen_entry(mt, theta, xm, kb, r2, c2, false));
    if (!deriv) {
      v = cmul(a, b);
    } else {
      const cd da = gen_entry(mt, theta, xm, kb, r1, c1, true);
      const cd db = cconj(gen_entry(mt, theta, xm, kb, r2, c2, true));
      v = cadd(cmul(da, b), cmul(a, db));
    }
  } else {
    v = gen_entry(mt, theta, xm, kb, r, c, deriv);
  }
  return (flags & JQ_MF_CONJ) ? cconj(v) : v;
}

__device__ __forceinline__ double member_theta(const int* mt, double cval, const double* angles) {
  const int p = mt[M_PARAM];
  return p >= 0 ? angles[p] : cval;
}

template <typename R>
struct Out2 {
  R x, y;
};

// Effective (active-block) matrix E of a member: D x D, or its diagonal.
__device__ void member_eff(const int* mt, double th, const cd* xm, int D, int offs, bool diagf, bool deriv, cd* E) {
  if (diagf) {
    for (int r = 0; r < D; ++r) E[r] = member_entry(mt, th, xm, offs + r, offs + r, deriv);
  } else {
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c) E[r * D + c] = member_entry(mt, th, xm, offs + r, offs + c, deriv);
  }
}

__device__ void matmul_d(const cd* A, const cd* B, cd* C, int D) {  // C = A B
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) {
      cd acc = {0.0, 0.0};
      for (int k = 0; k < D; ++k) acc = cadd(acc, cmul(A[r * D + k], B[k * D + c]));
      C[r * D + c] = acc;
    }
}

// Gauss-Jordan inverse with partial pivoting (D <= 8); returns false if singular.
__device__ bool invert_d(const cd* A, cd* Ainv, int D) {
  cd M[64], I[64];
  for (int e = 0; e < D * D; ++e) {
    M[e] = A[e];
    I[e] = {(e / D) == (e % D) ? 1.0 : 0.0, 0.0};
  }
  for (int col = 0; col < D; ++col) {
    int piv = col;
    double best = -1;
    for (int r = col; r < D; ++r) {
      const double m = M[r * D + col].x * M[r * D + col].x + M[r * D + col].y * M[r * D + col].y;
      if (m > best) best = m, piv = r;
    }
    if (!(best > 1e-300)) return false;
    if (piv != col)
      for (int c = 0; c < D; ++c) {
        cd t = M[col * D + c];
        M[col * D + c] = M[piv * D + c];
        M[piv * D + c] = t;
        t = I[col * D + c];
        I[col * D + c] = I[piv * D + c];
        I[piv * D + c] = t;
      }
    const cd p = M[col * D + col];
    const double den = p.x * p.x + p.y * p.y;
    const cd pinv = {p.x / den, -p.y / den};
    for (int c = 0; c < D; ++c) {
      M[col * D + c] = cmul(M[col * D + c], pinv);
      I[col * D + c] = cmul(I[col * D + c], pinv);
    }
    for (int r = 0; r < D; ++r) {
      if (r == col) continue;
      const cd f = M[r * D + col];
      if (f.x == 0.0 && f.y == 0.0) continue;
      for (int c = 0; c < D; ++c) {
        const cd a = cmul(f, M[col * D + c]);
        const cd b = cmul(f, I[col * D + c]);
        M[r * D + c] = {M[r * D + c].x - a.x, M[r * D + c].y - a.y};
        I[r * D + c] = {I[r * D + c].x - b.x, I[r * D + c].y - b.y};
      }
    }
  }
  for (int e = 0; e < D * D; ++e) Ainv[e] = I[e];
  return true;
}

// Effective matrix of a member with compile-time storage: DIAG keeps the
// diagonal (DM >= D entries used), dense blocks use D == DM (row stride DM),
// so every index is a constant and the arrays stay in registers.
template <bool DIAG, int DM>
__device__ __forceinline__ void member_eff_t(const int* mt, double th, const cd* xm, int D, int offs, bool deriv,
                                             cd* E) {
  if constexpr (DIAG) {
#pragma unroll
    for (int r = 0; r < DM; ++r)
      if (r < D) E[r] = member_entry(mt, th, xm, offs + r, offs + r, deriv);
  } else {
#pragma unroll
    for (int r = 0; r < DM; ++r)
#pragma unroll
      for (int c = 0; c < DM; ++c) E[r * DM + c] = member_entry(mt, th, xm, offs + r, offs + c, deriv);
  }
}

// M <- E M (dense, D == DM) or elementwise (diagonal)
template <bool DIAG, int DM>
__device__ __forceinline__ void left_mul_t(const cd* E, cd* M, int D) {
  if constexpr (DIAG) {
#pragma unroll
    for (int r = 0; r < DM; ++r)
      if (r < D) M[r] = cmul(E[r], M[r]);
  } else {
    cd T[DM * DM];
#pragma unroll
    for (int r = 0; r < DM; ++r)
#pragma unroll
      for (int c = 0; c < DM; ++c) {
        cd acc = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < DM; ++k) acc = cadd(acc, cmul(E[r * DM + k], M[k * DM + c]));
        T[r * DM + c] = acc;
      }
#pragma unroll
    for (int e = 0; e < DM * DM; ++e) M[e] = T[e];
  }
}

template <bool DIAG, int DM>
__device__ __forceinline__ void identity_t(cd* M) {
#pragma unroll
  for (int e = 0; e < (DIAG ? DM : DM * DM); ++e)
    M[e] = {DIAG ? 1.0 : ((e / DM) == (e % DM) ? 1.0 : 0.0), 0.0};
}

template <typename R, bool DIAG, int DM>
__device__ __forceinline__ void prep_block(const PrepArgs& a, int circ, const int* bt) {
  constexpr int N = DIAG ? DM : DM * DM;
  const int D = bt[B_D], offs = bt[B_OFFS];
  const int ne = DIAG ? D : D * D;
  const double* angles = a.angles + (int64_t)circ * a.n_params;
  const cd* xm = reinterpret_cast<const cd*>(a.matrices) + (a.matrices ? (int64_t)circ * a.n_matrix_elems : 0);
  cd M[N], E[N];
  identity_t<DIAG, DM>(M);
  for (int mi = bt[B_MBEG]; mi < bt[B_MEND]; ++mi) {
    const int m = a.member_list[mi];
    const int* mt = a.member_table + m * MW;
    const double th = member_theta(mt, a.member_values[m], angles);
    member_eff_t<DIAG, DM>(mt, th, xm, D, offs, false, E);
    left_mul_t<DIAG, DM>(E, M, D);
  }
  Out2<R>* pool = reinterpret_cast<Out2<R>*>(a.pool) + (int64_t)circ * a.pool_stride;
#pragma unroll
  for (int e = 0; e < N; ++e)
    if (e < ne) pool[bt[B_MAT] + e] = {(R)M[e].x, (R)M[e].y};
  if (bt[B_INV] >= 0) {
    cd T[N];
    bool ok = true;
    if constexpr (DIAG) {
#pragma unroll
      for (int r = 0; r < DM; ++r) {
        if (r >= D) continue;
        const double den = M[r].x * M[r].x + M[r].y * M[r].y;
        if (!(den > 1e-300)) ok = false;
        T[r] = {M[r].x / den, -M[r].y / den};
      }
    } else {
      ok = invert_d(M, T, DM);
    }
    (void)ok;  // the planner inverts only blocks of declared condition number <= JQ_COND_MAX
#pragma unroll
    for (int e = 0; e < N; ++e)
      if (e < ne) pool[bt[B_INV] + e] = {(R)T[e].x, (R)T[e].y};
  }
}

// One instance per block class (launch_prep starts those the plan holds):
// a kernel compiled for one matrix size keeps it in registers, and the
// scratch of the large classes is not reserved for the common small ones.
// (all three kernel families launch 64 threads per workgroup: the bound
// lets a thread keep its block's matrices in up to 512 VGPRs instead of
// scratch)
template <typename R, int CLS>
__global__ __launch_bounds__(64) void prep_kernel(PrepArgs a) {
  const int nb = a.n_blocks;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t)nb * a.batch) return;
  const int circ = (int)(gid / nb);
  const int b = (int)(gid % nb);
  const int* bt = a.block_table + b * BW;
  if (block_class(bt[B_DIAGF] != 0, bt[B_D]) != CLS) return;
  constexpr bool DIAG = CLS <= 1;
  constexpr int DM = CLS == 0 ? 4 : CLS == 1 ? 32 : 2;
  prep_block<R, DIAG, DM>(a, circ, bt);
}

// ---------------------------------------------------------------------------
// Workgroup forms for the larger blocks (4x4 and up): one workgroup of one
// wave per (block, circuit), one matrix entry per thread, matrices in LDS.
// A thread per block serialised their FP64 products (a 4x4 density-matrix
// block of nine members took ~1 ms per call that way).
// ---------------------------------------------------------------------------
constexpr int WG_NT = 64;
constexpr int WG_MAXE = 32 * 32;  // entries of the largest dense matrix

__device__ __forceinline__ void wg_identity(cd* M, int D, bool diagf) {
  const int ne = diagf ? D : D * D;
  for (int e = threadIdx.x; e < ne; e += WG_NT) M[e] = {diagf ? 1.0 : ((e / D) == (e % D) ? 1.0 : 0.0), 0.0};
}

__device__ __forceinline__ void wg_eff(const int* mt, double th, const cd* xm, int D, int offs, bool diagf, bool deriv,
                                       cd* E) {
  const int ne = diagf ? D : D * D;
  for (int e = threadIdx.x; e < ne; e += WG_NT) {
    const int r = diagf ? e : e / D, c = diagf ? e : e % D;
    E[e] = member_entry(mt, th, xm, offs + r, offs + c, deriv);
  }
}

// C = op(A) B, where op(A) = A or A^T (dense) or the elementwise product
// (diagonal); C aliases neither.
__device__ __forceinline__ void wg_mul(const cd* A, const cd* B, cd* C, int D, bool diagf, bool transA) {
  if (diagf) {
    for (int e = threadIdx.x; e < D; e += WG_NT) C[e] = cmul(A[e], B[e]);
    return;
  }
  for (int e = threadIdx.x; e < D * D; e += WG_NT) {
    const int r = e / D, c = e % D;
    cd acc = {0.0, 0.0};
    for (int k = 0; k < D; ++k) acc = cadd(acc, cmul(transA ? A[k * D + r] : A[r * D + k], B[k * D + c]));
    C[e] = acc;
  }
}

// C = A B^T (dense) or elementwise
__device__ __forceinline__ void wg_mul_bt(const cd* A, const cd* B, cd* C, int D, bool diagf) {
  if (diagf) {
    for (int e = threadIdx.x; e < D; e += WG_NT) C[e] = cmul(A[e], B[e]);
    return;
  }
  for (int e = threadIdx.x; e < D * D; e += WG_NT) {
    const int r = e / D, c = e % D;
    cd acc = {0.0, 0.0};
    for (int k = 0; k < D; ++k) acc = cadd(acc, cmul(A[r * D + k], B[c * D + k]));
    C[e] = acc;
  }
}

__device__ __forceinline__ void wg_copy(const cd* A, cd* B, int ne) {
  for (int e = threadIdx.x; e < ne; e += WG_NT) B[e] = A[e];
}

template <typename R>
__global__ __launch_bounds__(WG_NT) void prep_wg_kernel(PrepArgs a) {
  __shared__ cd M[WG_MAXE], E[WG_MAXE], T[WG_MAXE];
  const int b = a.wg_blocks[blockIdx.x];
  const int circ = blockIdx.y;
  const int* bt = a.block_table + b * BW;
  const int D = bt[B_D], offs = bt[B_OFFS];
  const bool diagf = bt[B_DIAGF] != 0;
  const int ne = diagf ? D : D * D;
  const double* angles = a.angles + (int64_t)circ * a.n_params;
  const cd* xm = reinterpret_cast<const cd*>(a.matrices) + (a.matrices ? (int64_t)circ * a.n_matrix_elems : 0);
  wg_identity(M, D, diagf);
  for (int mi = bt[B_MBEG]; mi < bt[B_MEND]; ++mi) {
    const int m = a.member_list[mi];
    const int* mt = a.member_table + m * MW;
    const double th = member_theta(mt, a.member_values[m], angles);
    __syncthreads();
    wg_eff(mt, th, xm, D, offs, diagf, false, E);
    __syncthreads();
    wg_mul(E, M, T, D, diagf, false);
    __syncthreads();
    wg_copy(T, M, ne);
  }
  __syncthreads();
  Out2<R>* pool = reinterpret_cast<Out2<R>*>(a.pool) + (int64_t)circ * a.pool_stride;
  for (int e = threadIdx.x; e < ne; e += WG_NT) pool[bt[B_MAT] + e] = {(R)M[e].x, (R)M[e].y};
  if (bt[B_INV] >= 0 && threadIdx.x == 0) {
    // tile-pass non-unitary blocks (<= 8 rows), of declared condition <= JQ_COND_MAX
    if (diagf) {
      for (int r = 0; r < D; ++r) {
        const double den = M[r].x * M[r].x + M[r].y * M[r].y;
        pool[bt[B_INV] + r] = {(R)(M[r].x / den), (R)(-M[r].y / den)};
      }
    } else if (D <= 8) {
      cd inv[64];
      invert_d(M, inv, D);
      for (int e = 0; e < D * D; ++e) pool[bt[B_INV] + e] = {(R)inv[e].x, (R)inv[e].y};
    }
  }
}

// sum over tiles of each cotangent slot -> bars (FP64 reals)
// One wave per output real: the per-tile partials of a real are contiguous
// ([real][tile] within each pass area), so the wave reads them coalesced and
// sums in a fixed order (bit-reproducible from run to run).
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double sum_row(const double* row, int n) {
  const int lane = threadIdx.x & 63;
  double s = 0;
  for (int t = lane; t < n; t += 64) s += row[t];
  return wave_sum_d(s);
}

__global__ void reduce_grad_kernel(ReduceArgs a) {
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wid >= (int64_t)a.n_reals * a.batch) return;
  const int circ = (int)(wid / a.n_reals);
  const int e = (int)(wid % a.n_reals);
  // find the slot-map entry holding real e of the cotangent array
  const int* sm = a.slot_map;
  int lo = 0, hi = a.n_slots - 1;
  while (lo < hi) {  // entries are sorted by bar offset
    const int mid = (lo + hi + 1) / 2;
    if (sm[mid * 5 + 3] <= e) lo = mid; else hi = mid - 1;
  }
  const int base = sm[lo * 5 + 0], local = sm[lo * 5 + 2], bar = sm[lo * 5 + 3];
  const double* row = a.partials + (int64_t)circ * a.part_stride + base + (int64_t)(local + (e - bar)) * a.ntiles;
  const double s = sum_row(row, a.ntiles);
  if ((threadIdx.x & 63) == 0) a.bars[(int64_t)circ * a.n_reals + e] = s;
}

// The same reduction with one thread per output real, summing its tiles in
// order: for circuits of few tiles a wave per output would mostly idle.
__global__ void reduce_grad_thread_kernel(ReduceArgs a) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)a.n_reals * a.batch) return;
  const int circ = (int)(id / a.n_reals);
  const int e = (int)(id % a.n_reals);
  const int* sm = a.slot_map;
  int lo = 0, hi = a.n_slots - 1;
  while (lo < hi) {  // entries are sorted by bar offset
    const int mid = (lo + hi + 1) / 2;
    if (sm[mid * 5 + 3] <= e) lo = mid; else hi = mid - 1;
  }
  const int base = sm[lo * 5 + 0], local = sm[lo * 5 + 2], bar = sm[lo * 5 + 3];
  const double* row = a.partials + (int64_t)circ * a.part_stride + base + (int64_t)(local + (e - bar)) * a.ntiles;
  double s = 0;
  for (int t = 0; t < a.ntiles; ++t) s += row[t];
  a.bars[(int64_t)circ * a.n_reals + e] = s;
}

// partials: [batch][n_obs][ntiles]
__global__ void reduce_ev_kernel(const double* part, double* out, int ntiles, int n_obs, int batch) {
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wid >= (int64_t)n_obs * batch) return;
  const double s = sum_row(part + wid * ntiles, ntiles);
  if ((threadIdx.x & 63) == 0) out[wid] = s;
}

// block cotangent -> member cotangents W_j = L_j^T bar R_j^T -> Re sum W . dE
template <bool DIAG, int DM>
__device__ __forceinline__ void chain_block(const ChainArgs& a, int circ, const int* bt) {
  constexpr int N = DIAG ? DM : DM * DM;
  const int D = bt[B_D], offs = bt[B_OFFS];
  const int ne = DIAG ? D : D * D;
  const double* angles = a.angles + (int64_t)circ * a.n_params;
  const cd* xm = reinterpret_cast<const cd*>(a.matrices) + (a.matrices ? (int64_t)circ * a.n_matrix_elems : 0);
  const cd* bar = reinterpret_cast<const cd*>(a.bars) + (int64_t)circ * a.bars_stride + bt[B_BAR];
  double* gang = a.grad_angles ? a.grad_angles + (int64_t)circ * a.n_params : nullptr;
  cd* gmat = a.grad_matrices ? reinterpret_cast<cd*>(a.grad_matrices) + (int64_t)circ * a.n_matrix_elems : nullptr;
  const int mb = bt[B_MBEG], me = bt[B_MEND];
  cd Rp[N], Lm[N], E[N], W[N], B[N];
#pragma unroll
  for (int e = 0; e < N; ++e) B[e] = e < ne ? bar[e] : cd{0.0, 0.0};
  if constexpr (!DIAG && DM == 2) {
    if (a.quat_t && bt[B_QUAT]) {
      // The block U = [[al, -conj(be)], [be, conj(al)]] is in SU(2) and every
      // angle derivative is U X with X in su(2): only t_k = Im Tr(bar^T U
      // sigma_k) enters the gradients.  The effective cotangent with zero
      // bar01 / bar11 giving the same Re Tr(bar^T U X) for every X is
      //   bar00 = -be (i t_x + t_y) + i t_z conj(al),
      //   bar10 =  al (i t_x + t_y) + i t_z conj(be).
      identity_t<DIAG, DM>(Lm);
      for (int j = mb; j < me; ++j) {
        const int m = a.member_list[j];
        const int* mt = a.member_table + m * MW;
        member_eff_t<DIAG, DM>(mt, member_theta(mt, a.member_values[m], angles), xm, D, offs, false, E);
        left_mul_t<DIAG, DM>(E, Lm, D);
      }
      const cd al = Lm[0], be = Lm[2];
      const double tx = B[0].x, ty = B[0].y, tz = B[2].x;
      const cd w = {ty, tx};  // i t_x + t_y
      const cd izt = {0.0, tz};
      const cd mbe = {-be.x, -be.y};
      B[0] = cadd(cmul(mbe, w), cmul(izt, cd{al.x, -al.y}));
      B[2] = cadd(cmul(al, w), cmul(izt, cd{be.x, -be.y}));
      B[1] = cd{0.0, 0.0};
      B[3] = cd{0.0, 0.0};
    }
  }
  identity_t<DIAG, DM>(Rp);
  for (int j = mb; j < me; ++j) {
    const int m = a.member_list[j];
    const int* mt = a.member_table + m * MW;
    const double th = member_theta(mt, a.member_values[m], angles);
    const bool want_theta = mt[M_PARAM] >= 0 && gang != nullptr;
    const bool want_mat = mt[M_KIND] == JQ_M_MATRIX && (mt[M_FLAGS] & JQ_MF_GRAD) && gmat != nullptr;
    if (want_theta || want_mat) {
      // L = E_last ... E_{j+1}
      identity_t<DIAG, DM>(Lm);
      for (int q = j + 1; q < me; ++q) {
        const int mq = a.member_list[q];
        const int* mtq = a.member_table + mq * MW;
        member_eff_t<DIAG, DM>(mtq, member_theta(mtq, a.member_values[mq], angles), xm, D, offs, false, E);
        left_mul_t<DIAG, DM>(E, Lm, D);
      }
      // W = L^T bar R^T
      if constexpr (DIAG) {
#pragma unroll
        for (int r = 0; r < DM; ++r)
          if (r < D) W[r] = cmul(cmul(Lm[r], B[r]), Rp[r]);
      } else {
        cd T[N];
#pragma unroll
        for (int r = 0; r < DM; ++r)
#pragma unroll
          for (int c = 0; c < DM; ++c) {
            cd acc = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < DM; ++k) acc = cadd(acc, cmul(Lm[k * DM + r], B[k * DM + c]));
            T[r * DM + c] = acc;
          }
#pragma unroll
        for (int r = 0; r < DM; ++r)
#pragma unroll
          for (int c = 0; c < DM; ++c) {
            cd acc = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < DM; ++k) acc = cadd(acc, cmul(T[r * DM + k], Rp[c * DM + k]));
            W[r * DM + c] = acc;
          }
      }
      if (want_theta) {
        member_eff_t<DIAG, DM>(mt, th, xm, D, offs, true, E);
        double g = 0;
#pragma unroll
        for (int e = 0; e < N; ++e)
          if (e < ne) g += W[e].x * E[e].x - W[e].y * E[e].y;
        atomicAdd(&gang[mt[M_PARAM]], g);
      }
      if (want_mat) {
        const int dim = 1 << mt[M_KB];
        // explicit matrices are unstructured: D == dim, offs == 0
#pragma unroll
        for (int r = 0; r < DM; ++r)
#pragma unroll
          for (int c = 0; c < DM; ++c) {
            if (r >= D || c >= D) continue;
            const cd w = DIAG ? (r == c ? W[r] : cd{0.0, 0.0}) : W[r * DM + c];
            cd* dst = &gmat[mt[M_MATRIX] + r * dim + c];
            atomicAdd(&dst->x, w.x);
            atomicAdd(&dst->y, w.y);
          }
      }
    }
    // R <- E_j R
    member_eff_t<DIAG, DM>(mt, th, xm, D, offs, false, E);
    left_mul_t<DIAG, DM>(E, Rp, D);
  }
}

// The 2x2 class with one thread per (circuit, block, member j): the thread
// rebuilds only its own prefix R_j = E_{j-1} .. E_0 and suffix L_j = E_last
// .. E_{j+1} (j and m-1-j factors) instead of one thread walking every
// member of the block and every suffix in turn, so the chain's serial depth
// is one member product instead of the block's whole quadratic walk.
// Rotation blocks in su(2) form (quat_t) take U = L_j E_j R_j.
__global__ __launch_bounds__(64) void chain2_kernel(ChainArgs a) {
  constexpr int DM = 2, N = 4;
  const int mm = a.max_members2;
  const int nb = a.n_blocks;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t)nb * a.batch * mm) return;
  const int jo = (int)(gid % mm);
  const int64_t bc = gid / mm;
  const int circ = (int)(bc / nb);
  const int b = (int)(bc % nb);
  const int* bt = a.block_table + b * BW;
  if (bt[B_BAR] < 0 || block_class(bt[B_DIAGF] != 0, bt[B_D]) != 2) return;
  const int mb = bt[B_MBEG], me = bt[B_MEND];
  const int j = mb + jo;
  if (j >= me) return;
  const int D = bt[B_D], offs = bt[B_OFFS];
  const double* angles = a.angles + (int64_t)circ * a.n_params;
  const cd* xm = reinterpret_cast<const cd*>(a.matrices) + (a.matrices ? (int64_t)circ * a.n_matrix_elems : 0);
  double* gang = a.grad_angles ? a.grad_angles + (int64_t)circ * a.n_params : nullptr;
  cd* gmat = a.grad_matrices ? reinterpret_cast<cd*>(a.grad_matrices) + (int64_t)circ * a.n_matrix_elems : nullptr;
  const int m = a.member_list[j];
  const int* mt = a.member_table + m * MW;
  const bool want_theta = mt[M_PARAM] >= 0 && gang != nullptr;
  const bool want_mat = mt[M_KIND] == JQ_M_MATRIX && (mt[M_FLAGS] & JQ_MF_GRAD) && gmat != nullptr;
  if (!want_theta && !want_mat) return;
  const cd* bar = reinterpret_cast<const cd*>(a.bars) + (int64_t)circ * a.bars_stride + bt[B_BAR];
  cd Rp[N], Lm[N], E[N], W[N], B[N];
#pragma unroll
  for (int e = 0; e < N; ++e) B[e] = bar[e];
  identity_t<false, DM>(Rp);
  for (int q = mb; q < j; ++q) {
    const int* mtq = a.member_table + a.member_list[q] * MW;
    member_eff_t<false, DM>(mtq, member_theta(mtq, a.member_values[a.member_list[q]], angles), xm, D, offs, false, E);
    left_mul_t<false, DM>(E, Rp, D);
  }
  identity_t<false, DM>(Lm);
  for (int q = j + 1; q < me; ++q) {
    const int* mtq = a.member_table + a.member_list[q] * MW;
    member_eff_t<false, DM>(mtq, member_theta(mtq, a.member_values[a.member_list[q]], angles), xm, D, offs, false, E);
    left_mul_t<false, DM>(E, Lm, D);
  }
  const double th = member_theta(mt, a.member_values[m], angles);
  if (a.quat_t && bt[B_QUAT]) {
    // effective cotangent of U from (t_x, t_y, t_z), as in chain_block
    cd U[N];
#pragma unroll
    for (int e = 0; e < N; ++e) U[e] = Rp[e];
    member_eff_t<false, DM>(mt, th, xm, D, offs, false, E);
    left_mul_t<false, DM>(E, U, D);
    left_mul_t<false, DM>(Lm, U, D);
    const cd al = U[0], be = U[2];
    const double tx = B[0].x, ty = B[0].y, tz = B[2].x;
    const cd w = {ty, tx};
    const cd izt = {0.0, tz};
    const cd mbe = {-be.x, -be.y};
    B[0] = cadd(cmul(mbe, w), cmul(izt, cd{al.x, -al.y}));
    B[2] = cadd(cmul(al, w), cmul(izt, cd{be.x, -be.y}));
    B[1] = cd{0.0, 0.0};
    B[3] = cd{0.0, 0.0};
  }
  // W = L^T bar R^T
  {
    cd T[N];
#pragma unroll
    for (int r = 0; r < DM; ++r)
#pragma unroll
      for (int c = 0; c < DM; ++c) {
        cd acc = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < DM; ++k) acc = cadd(acc, cmul(Lm[k * DM + r], B[k * DM + c]));
        T[r * DM + c] = acc;
      }
#pragma unroll
    for (int r = 0; r < DM; ++r)
#pragma unroll
      for (int c = 0; c < DM; ++c) {
        cd acc = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < DM; ++k) acc = cadd(acc, cmul(T[r * DM + k], Rp[c * DM + k]));
        W[r * DM + c] = acc;
      }
  }
  if (want_theta) {
    member_eff_t<false, DM>(mt, th, xm, D, offs, true, E);
    double g = 0;
#pragma unroll
    for (int e = 0; e < N; ++e) g += W[e].x * E[e].x - W[e].y * E[e].y;
    atomicAdd(&gang[mt[M_PARAM]], g);
  }
  if (want_mat) {
    const int dim = 1 << mt[M_KB];
#pragma unroll
    for (int r = 0; r < DM; ++r)
#pragma unroll
      for (int c = 0; c < DM; ++c) {
        cd* dst = &gmat[mt[M_MATRIX] + r * dim + c];
        atomicAdd(&dst->x, W[r * DM + c].x);
        atomicAdd(&dst->y, W[r * DM + c].y);
      }
  }
}

template <int CLS>
__global__ __launch_bounds__(64) void chain_kernel(ChainArgs a) {
  const int nb = a.n_blocks;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t)nb * a.batch) return;
  const int circ = (int)(gid / nb);
  const int b = (int)(gid % nb);
  const int* bt = a.block_table + b * BW;
  if (bt[B_BAR] < 0 || block_class(bt[B_DIAGF] != 0, bt[B_D]) != CLS) return;
  constexpr bool DIAG = CLS <= 1;
  constexpr int DM = CLS == 0 ? 4 : CLS == 1 ? 32 : 2;
  chain_block<DIAG, DM>(a, circ, bt);
}

// Workgroup form of chain_block: W_j = L_j^T bar R_j^T with the prefix R_j
// kept in LDS and the suffix L_j rebuilt per differentiated member.
__global__ __launch_bounds__(WG_NT) void chain_wg_kernel(ChainArgs a) {
  __shared__ cd Rp[WG_MAXE], Lm[WG_MAXE], E[WG_MAXE], T[WG_MAXE], B[WG_MAXE];
  __shared__ double red[WG_NT / 64];
  const int b = a.wg_blocks[blockIdx.x];
  const int circ = blockIdx.y;
  const int* bt = a.block_table + b * BW;
  if (bt[B_BAR] < 0) return;
  const int D = bt[B_D], offs = bt[B_OFFS];
  const bool diagf = bt[B_DIAGF] != 0;
  const int ne = diagf ? D : D * D;
  const double* angles = a.angles + (int64_t)circ * a.n_params;
  const cd* xm = reinterpret_cast<const cd*>(a.matrices) + (a.matrices ? (int64_t)circ * a.n_matrix_elems : 0);
  const cd* bar = reinterpret_cast<const cd*>(a.bars) + (int64_t)circ * a.bars_stride + bt[B_BAR];
  double* gang = a.grad_angles ? a.grad_angles + (int64_t)circ * a.n_params : nullptr;
  cd* gmat = a.grad_matrices ? reinterpret_cast<cd*>(a.grad_matrices) + (int64_t)circ * a.n_matrix_elems : nullptr;
  const int mb = bt[B_MBEG], me = bt[B_MEND];
  wg_copy(bar, B, ne);
  wg_identity(Rp, D, diagf);
  for (int j = mb; j < me; ++j) {
    const int m = a.member_list[j];
    const int* mt = a.member_table + m * MW;
    const double th = member_theta(mt, a.member_values[m], angles);
    const bool want_theta = mt[M_PARAM] >= 0 && gang != nullptr;
    const bool want_mat = mt[M_KIND] == JQ_M_MATRIX && (mt[M_FLAGS] & JQ_MF_GRAD) && gmat != nullptr;
    __syncthreads();
    if (want_theta || want_mat) {
      wg_identity(Lm, D, diagf);
      for (int q = j + 1; q < me; ++q) {
        const int mq = a.member_list[q];
        const int* mtq = a.member_table + mq * MW;
        __syncthreads();
        wg_eff(mtq, member_theta(mtq, a.member_values[mq], angles), xm, D, offs, diagf, false, E);
        __syncthreads();
        wg_mul(E, Lm, T, D, diagf, false);
        __syncthreads();
        wg_copy(T, Lm, ne);
      }
      __syncthreads();
      wg_mul(Lm, B, T, D, diagf, true);  // L^T bar
      __syncthreads();
      wg_mul_bt(T, Rp, Lm, D, diagf);    // W = L^T bar R^T (into Lm)
      __syncthreads();
      if (want_theta) {
        wg_eff(mt, th, xm, D, offs, diagf, true, E);
        __syncthreads();
        double g = 0;
        for (int e = threadIdx.x; e < ne; e += WG_NT) g += Lm[e].x * E[e].x - Lm[e].y * E[e].y;
        g = wave_sum_d(g);
        if (threadIdx.x == 0) atomicAdd(&gang[mt[M_PARAM]], g);
      }
      if (want_mat) {
        const int dim = 1 << mt[M_KB];  // explicit matrices are unstructured: D == dim, offs == 0
        for (int e = threadIdx.x; e < ne; e += WG_NT) {
          const int r = diagf ? e : e / D, c = diagf ? e : e % D;
          cd* dst = &gmat[mt[M_MATRIX] + r * dim + c];
          atomicAdd(&dst->x, Lm[e].x);
          atomicAdd(&dst->y, Lm[e].y);
        }
      }
      __syncthreads();
    }
    wg_eff(mt, th, xm, D, offs, diagf, false, E);
    __syncthreads();
    wg_mul(E, Rp, T, D, diagf, false);
    __syncthreads();
    wg_copy(T, Rp, ne);
  }
  (void)red;
}

// ---------------------------------------------------------------------------
// measurement outputs
// ---------------------------------------------------------------------------
template <typename R>
__global__ void probs_kernel(const Out2<R>* s, R* p, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const Out2<R> v = s[i];
  p[i] = v.x * v.x + v.y * v.y;
}

template <typename R>
__global__ void outer_kernel(const Out2<R>* s, Out2<R>* rho, int64_t dim, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t b = i / (dim * dim), rc = i % (dim * dim);
  const Out2<R> x = s[b * dim + rc / dim], y = s[b * dim + rc % dim];
  rho[i] = {x.x * y.x + x.y * y.y, x.y * y.x - x.x * y.y};  // x conj(y)
}

template <typename R>
__global__ void dm_diag_kernel(const Out2<R>* rho, R* p, int64_t dim, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t b = i / dim, r = i % dim;
  p[i] = rho[b * dim * dim + r * dim + r].x;
}

// one workgroup per (state, chunk): sum over groups of 2^k amplitudes of
// conj(psi_g)^T O psi_g, or Tr(O rho) over the ket wires of rho.
template <typename R>
__global__ void expval_matrix_kernel(ExpvalArgs a) {
  const int circ = blockIdx.y;
  const int k = a.k;
  const int D = 1 << k;
  const int64_t ngroups = (int64_t)1 << (a.n_qubits - k);
  const Out2<R>* s = reinterpret_cast<const Out2<R>*>(a.state) + (int64_t)circ * a.stride;
  const int64_t dim = (int64_t)1 << a.n_qubits;
  const cd* om = reinterpret_cast<const cd*>(a.mat);
  double acc = 0;
  for (int64_t grp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; grp < ngroups;
       grp += (int64_t)gridDim.x * blockDim.x) {
    // deposit the group index into the bits of the other wires
    int64_t idx0 = 0;
    int src = 0;
    for (int gb = 0; gb < a.n_qubits; ++gb) {
      if (a.wmask & (1ll << gb)) continue;
      if ((grp >> src) & 1) idx0 |= 1ll << gb;
      ++src;
    }
    int64_t idx[32];
    for (int q = 0; q < D; ++q) {
      int64_t x = idx0;
      for (int j = 0; j < k; ++j)
        if ((q >> (k - 1 - j)) & 1) x |= 1ll << a.gbits[j];
      idx[q] = x;
    }
    if (!a.density) {
      cd x[32];
      for (int q = 0; q < D; ++q) x[q] = {(double)s[idx[q]].x, (double)s[idx[q]].y};
      for (int r = 0; r < D; ++r) {
        cd y = {0, 0};
        for (int c = 0; c < D; ++c) y = cadd(y, cmul(om[r * D + c], x[c]));
        acc += x[r].x * y.x + x[r].y * y.y;  // Re conj(x_r) y_r
      }
    } else {
      // Tr(O rho) restricted to this rest index: sum_{r,c} O[r][c] rho[c][r]
      for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
          const Out2<R> v = s[idx[c] * dim + idx[r]];
          const cd o = om[r * D + c];
          acc += o.x * (double)v.x - o.y * (double)v.y;
        }
    }
  }
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(&a.out[circ], acc);
}

__global__ void obs_upload_kernel(ObsTable t, ObsTable* dst) {
  const int i = threadIdx.x;
  if (i < 64) dst->gbit[i] = t.gbit[i];
  if (i < 128) dst->d[i] = t.d[i];
  if (i < 64) dst->mask[i] = t.mask[i];
}

hipError_t launch_obs_upload(const ObsTable& t, void* dst, hipStream_t st) {
  hipLaunchKernelGGL(obs_upload_kernel, dim3(1), dim3(128), 0, st, t, (ObsTable*)dst);
  return hipGetLastError();
}

struct ByteChunk {
  uint32_t w[512];  // 2 KiB per launch, passed by value
};

__global__ void bytes_upload_kernel(ByteChunk c, uint32_t nwords, uint32_t* dst) {
  for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = c.w[i];
}

hipError_t launch_bytes_upload(const void* src, size_t n, void* dst, hipStream_t st) {
  if (n % 4 != 0) return hipErrorInvalidValue;
  const uint32_t* s = (const uint32_t*)src;
  uint32_t* d = (uint32_t*)dst;
  size_t words = n / 4;
  while (words > 0) {
    ByteChunk c;
    const uint32_t k = (uint32_t)(words < 512 ? words : 512);
    for (uint32_t i = 0; i < k; ++i) c.w[i] = s[i];
    hipLaunchKernelGGL(bytes_upload_kernel, dim3(1), dim3(256), 0, st, c, k, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    s += k;
    d += k;
    words -= k;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------
hipError_t launch_prep(int dtype, const PrepArgs& a, hipStream_t st) {
  const int64_t n = (int64_t)a.n_blocks * a.batch;
  if (n == 0) return hipSuccess;
  const int tpb = 64;
  dim3 grid((unsigned)((n + tpb - 1) / tpb));
#define JQ_PREP(C)                                                                      \
  if (a.class_mask & (1 << C)) {                                                      \
    if (dtype == 0)                                                                   \
      hipLaunchKernelGGL((prep_kernel<float, C>), grid, dim3(tpb), 0, st, a);         \
    else                                                                              \
      hipLaunchKernelGGL((prep_kernel<double, C>), grid, dim3(tpb), 0, st, a);        \
  }
  JQ_PREP(0) JQ_PREP(1) JQ_PREP(2)
#undef JQ_PREP
  if (a.n_wg > 0) {
    const dim3 g((unsigned)a.n_wg, (unsigned)a.batch);
    if (dtype == 0)
      hipLaunchKernelGGL(prep_wg_kernel<float>, g, dim3(WG_NT), 0, st, a);
    else
      hipLaunchKernelGGL(prep_wg_kernel<double>, g, dim3(WG_NT), 0, st, a);
  }
  return hipGetLastError();
}

hipError_t launch_reduce_grad(const ReduceArgs& a, hipStream_t st) {
  const int64_t n = (int64_t)a.n_reals * a.batch;
  if (n == 0) return hipSuccess;
  if (a.ntiles < 64)  // few tiles per circuit (many small circuits): one thread per output
    hipLaunchKernelGGL(reduce_grad_thread_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(reduce_grad_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Zero fill by kernel rather than hipMemsetAsync: the entry points stay
// safe to record into a HIP graph (memset nodes have given wrong results on
// replays after the first).
__global__ void zero_words_kernel(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  if (bytes % 4 != 0 || ((uintptr_t)p & 3u) != 0) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(bytes / 4);
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, n);
  return hipGetLastError();
}

hipError_t launch_reduce_ev(const double* part, double* out, int ntiles, int n_obs, int batch, hipStream_t st) {
  const int64_t n = (int64_t)n_obs * batch;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_ev_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, part, out, ntiles, n_obs,
                     batch);
  return hipGetLastError();
}

hipError_t launch_chain(const ChainArgs& a, hipStream_t st) {
  const int64_t n = (int64_t)a.n_blocks * a.batch;
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 63) / 64));
  if (a.class_mask & 1) hipLaunchKernelGGL(chain_kernel<0>, grid, dim3(64), 0, st, a);
  if (a.class_mask & 2) hipLaunchKernelGGL(chain_kernel<1>, grid, dim3(64), 0, st, a);
  if (a.class_mask & 4) {
    if (a.max_members2 > 0) {
      const int64_t n2 = n * a.max_members2;
      hipLaunchKernelGGL(chain2_kernel, dim3((unsigned)((n2 + 63) / 64)), dim3(64), 0, st, a);
    } else {
      hipLaunchKernelGGL(chain_kernel<2>, grid, dim3(64), 0, st, a);
    }
  }
  if (a.n_wg > 0) hipLaunchKernelGGL(chain_wg_kernel, dim3((unsigned)a.n_wg, (unsigned)a.batch), dim3(WG_NT), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_probs(int dtype, const void* s, void* p, int64_t total, hipStream_t st) {
  dim3 g((unsigned)((total + 255) / 256));
  if (dtype == 0)
    hipLaunchKernelGGL(probs_kernel<float>, g, dim3(256), 0, st, (const Out2<float>*)s, (float*)p, total);
  else
    hipLaunchKernelGGL(probs_kernel<double>, g, dim3(256), 0, st, (const Out2<double>*)s, (double*)p, total);
  return hipGetLastError();
}

hipError_t launch_outer(int dtype, const void* s, void* rho, int64_t dim, int64_t batch, hipStream_t st) {
  const int64_t total = dim * dim * batch;
  dim3 g((unsigned)((total + 255) / 256));
  if (dtype == 0)
    hipLaunchKernelGGL(outer_kernel<float>, g, dim3(256), 0, st, (const Out2<float>*)s, (Out2<float>*)rho, dim, total);
  else
    hipLaunchKernelGGL(outer_kernel<double>, g, dim3(256), 0, st, (const Out2<double>*)s, (Out2<double>*)rho, dim,
                       total);
  return hipGetLastError();
}

hipError_t launch_dm_diag(int dtype, const void* rho, void* p, int64_t dim, int64_t batch, hipStream_t st) {
  const int64_t total = dim * batch;
  dim3 g((unsigned)((total + 255) / 256));
  if (dtype == 0)
    hipLaunchKernelGGL(dm_diag_kernel<float>, g, dim3(256), 0, st, (const Out2<float>*)rho, (float*)p, dim, total);
  else
    hipLaunchKernelGGL(dm_diag_kernel<double>, g, dim3(256), 0, st, (const Out2<double>*)rho, (double*)p, dim, total);
  return hipGetLastError();
}

hipError_t launch_expval_matrix(int dtype, const ExpvalArgs& a, int batch, hipStream_t st) {
  const int64_t ngroups = (int64_t)1 << (a.n_qubits - a.k);
  int64_t blocks = (ngroups + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  dim3 g((unsigned)blocks, (unsigned)batch);
  if (dtype == 0)
    hipLaunchKernelGGL(expval_matrix_kernel<float>, g, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(expval_matrix_kernel<double>, g, dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace jq

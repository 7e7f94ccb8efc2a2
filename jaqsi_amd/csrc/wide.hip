// Wide-block passes: one dense (<= 5 wires) or diagonal (<= 6 wires) block
// applied to the whole vector, and its adjoint.
//
// The reference applies a k-qubit matrix by one einsum over the state
// (jaqsi/simulation.py:92-94) and gets its cotangent from the custom_vjp
// (simulation.py:616-650: bar[o][i] = sum_rest lambda[o, rest] psi[i, rest],
// lambda <- M^T lambda); for channels it never inverts anything
// (simulation.py:540-543, 368-381).  Here a workgroup owns a contiguous
// range of "rest" indices (all wire bits zero) and walks it in sub-chunks
// held in LDS: the 2^k amplitudes of each rest index are gathered column by
// column (consecutive threads take consecutive rest indices, so the global
// accesses stay coalesced), multiplied by the block matrix from LDS, and
// written back in place.  The backward reads the pass's input state from its
// checkpoint, accumulates the cotangent of the block per workgroup in FP64
// (the same per-tile partial layout the tile passes use, so one reduction
// and chain-rule kernel serve both) and updates lambda in place.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "jaqsi_amd.h"
#include "program.h"
#include "wide.h"

namespace jq {

namespace {

template <typename R>
struct V2 {
  R x, y;
};

constexpr int NT = 256;

__device__ __forceinline__ uint64_t expand_rest(uint64_t r, const WideGeom& g) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= g.k) break;
    const int p = g.sorted[j];
    r = ((r >> p) << (p + 1)) | (r & ((1ull << p) - 1));
  }
  return r;
}

__device__ __forceinline__ uint64_t wire_bits(int c, const WideGeom& g) {
  uint64_t x = 0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (j >= g.k) break;
    x |= (uint64_t)((c >> (g.k - 1 - j)) & 1) << g.gb[j];
  }
  return x;
}

template <typename R>
__device__ __forceinline__ V2<R> read_amp(const V2<R>* src, uint64_t idx) {
  if (src != nullptr) return src[idx];
  V2<R> v;
  v.x = idx == 0 ? R(1) : R(0);
  v.y = R(0);
  return v;
}

template <typename R>
__global__ __launch_bounds__(NT) void wide_fwd_kernel(SvArgs a, WideGeom g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int circ = blockIdx.y;
  const V2<R>* src = a.src ? reinterpret_cast<const V2<R>*>(a.src) + (int64_t)circ * a.src_stride : nullptr;
  V2<R>* dst = reinterpret_cast<V2<R>*>(a.dst) + (int64_t)circ * a.dst_stride;
  const V2<R>* M = reinterpret_cast<const V2<R>*>(a.mats) + (int64_t)circ * a.pool_stride + g.mat;
  const int D = g.D, lgD = g.lgD, lgS = g.lgS;
  const int S = 1 << lgS, SP = S + 1;
  V2<R>* X = reinterpret_cast<V2<R>*>(smem);
  V2<R>* Y = X + D * SP;
  V2<R>* Ms = Y + D * SP;
  if (!g.diag)
    for (int e = threadIdx.x; e < D * D; e += NT) Ms[e] = M[e];
  const uint64_t r_begin = (uint64_t)blockIdx.x * g.rpw;
  for (int64_t r0 = 0; r0 < g.rpw; r0 += S) {
    if (g.diag) {  // elementwise: no staging
      for (int e = threadIdx.x; e < (S << lgD); e += NT) {
        const int c = e >> lgS, rr = e & (S - 1);
        const uint64_t idx = expand_rest(r_begin + r0 + rr, g) | wire_bits(c, g);
        const V2<R> v = read_amp(src, idx), d = M[c];
        V2<R> o;
        o.x = d.x * v.x - d.y * v.y;
        o.y = d.x * v.y + d.y * v.x;
        dst[idx] = o;
      }
      continue;
    }
    __syncthreads();  // Ms loaded; the previous sub-chunk's Y stored
    for (int e = threadIdx.x; e < (S << lgD); e += NT) {
      const int c = e >> lgS, rr = e & (S - 1);
      X[c * SP + rr] = read_amp(src, expand_rest(r_begin + r0 + rr, g) | wire_bits(c, g));
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (S << lgD); e += NT) {
      const int o = e >> lgS, rr = e & (S - 1);
      R ar = 0, ai = 0;
      for (int c = 0; c < D; ++c) {
        const V2<R> m = Ms[o * D + c], x = X[c * SP + rr];
        ar += m.x * x.x - m.y * x.y;
        ai += m.x * x.y + m.y * x.x;
      }
      Y[o * SP + rr] = V2<R>{ar, ai};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (S << lgD); e += NT) {
      const int c = e >> lgS, rr = e & (S - 1);
      dst[expand_rest(r_begin + r0 + rr, g) | wire_bits(c, g)] = Y[c * SP + rr];
    }
  }
}

// at most 1024 cotangent entries (32 x 32) over 256 threads
constexpr int MAX_PER_THREAD = 4;

template <typename R>
__global__ __launch_bounds__(NT) void wide_bwd_kernel(SvArgs a, WideGeom g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int circ = blockIdx.y;
  const V2<R>* src = a.src ? reinterpret_cast<const V2<R>*>(a.src) + (int64_t)circ * a.src_stride : nullptr;
  V2<R>* lam = reinterpret_cast<V2<R>*>(a.lam) + (int64_t)circ * a.lam_stride;
  const V2<R>* M = reinterpret_cast<const V2<R>*>(a.mats) + (int64_t)circ * a.pool_stride + g.mat;
  const int D = g.D, lgD = g.lgD, lgS = g.lgS;
  const int S = 1 << lgS, SP = S + 1;
  const int NB = g.diag ? D : D * D;  // cotangent entries
  const bool grad = g.slot >= 0;
  V2<R>* X = reinterpret_cast<V2<R>*>(smem);  // psi_in  [c][rr]
  V2<R>* L = X + D * SP;                       // lambda_out [o][rr]
  V2<R>* Y = L + D * SP;                       // lambda_in  [c][rr]
  V2<R>* Ms = Y + D * SP;
  if (!g.diag)
    for (int e = threadIdx.x; e < D * D; e += NT) Ms[e] = M[e];
  double accr[MAX_PER_THREAD], acci[MAX_PER_THREAD];
#pragma unroll
  for (int j = 0; j < MAX_PER_THREAD; ++j) accr[j] = acci[j] = 0.0;
  const uint64_t r_begin = (uint64_t)blockIdx.x * g.rpw;
  for (int64_t r0 = 0; r0 < g.rpw; r0 += S) {
    __syncthreads();  // Ms loaded; the previous sub-chunk consumed
    for (int e = threadIdx.x; e < (S << lgD); e += NT) {
      const int c = e >> lgS, rr = e & (S - 1);
      const uint64_t idx = expand_rest(r_begin + r0 + rr, g) | wire_bits(c, g);
      X[c * SP + rr] = read_amp(src, idx);
      L[c * SP + rr] = lam[idx];
    }
    __syncthreads();
    if (grad) {
#pragma unroll
      for (int j = 0; j < MAX_PER_THREAD; ++j) {
        const int e = threadIdx.x + j * NT;
        if (e >= NB) break;
        const int o = g.diag ? e : (e >> lgD), c = g.diag ? e : (e & (D - 1));
        double sr = 0, si = 0;
        for (int rr = 0; rr < S; ++rr) {
          const V2<R> l = L[o * SP + rr], x = X[c * SP + rr];
          sr += (double)(l.x * x.x - l.y * x.y);
          si += (double)(l.x * x.y + l.y * x.x);
        }
        accr[j] += sr;
        acci[j] += si;
      }
    }
    // lambda_in = M^T lambda_out
    for (int e = threadIdx.x; e < (S << lgD); e += NT) {
      const int c = e >> lgS, rr = e & (S - 1);
      R ar = 0, ai = 0;
      if (g.diag) {
        const V2<R> d = M[c], l = L[c * SP + rr];
        ar = d.x * l.x - d.y * l.y;
        ai = d.x * l.y + d.y * l.x;
      } else {
        for (int o = 0; o < D; ++o) {
          const V2<R> m = Ms[o * D + c], l = L[o * SP + rr];
          ar += m.x * l.x - m.y * l.y;
          ai += m.x * l.y + m.y * l.x;
        }
      }
      Y[c * SP + rr] = V2<R>{ar, ai};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (S << lgD); e += NT) {
      const int c = e >> lgS, rr = e & (S - 1);
      lam[expand_rest(r_begin + r0 + rr, g) | wire_bits(c, g)] = Y[c * SP + rr];
    }
  }
  if (grad) {
    double* part = a.grad_part + (int64_t)circ * a.part_stride + g.slot_base;
#pragma unroll
    for (int j = 0; j < MAX_PER_THREAD; ++j) {
      const int e = threadIdx.x + j * NT;
      if (e >= NB) break;
      part[(int64_t)(g.slot + 2 * e) * g.ntiles + blockIdx.x] = accr[j];
      part[(int64_t)(g.slot + 2 * e + 1) * g.ntiles + blockIdx.x] = acci[j];
    }
  }
}

size_t smem_bytes(int dir, int dtype, const WideGeom& g) {
  const size_t eb = dtype == 0 ? 8 : 16;
  const size_t S = (size_t)1 << g.lgS;
  const size_t bufs = dir == 0 ? 2 : 3;
  return eb * (bufs * g.D * (S + 1) + (g.diag ? 0 : (size_t)g.D * g.D));
}

}  // namespace

void wide_geom(const int32_t* gw, int n_wires, int ntiles, int slot_base, int dtype, WideGeom& g) {
  g.k = gw[G_K];
  g.D = 1 << g.k;
  g.lgD = g.k;
  g.diag = gw[G_RA] != 0;
  for (int j = 0; j < 6; ++j) g.gb[j] = g.sorted[j] = 0;
  for (int j = 0; j < g.k; ++j) g.gb[j] = g.sorted[j] = (gw[G_DW0 + j / 4] >> (8 * (j % 4))) & 0xff;
  std::sort(g.sorted, g.sorted + g.k);
  g.rpw = ((int64_t)1 << (n_wires - g.k)) / ntiles;
  // sub-chunks of at most 4096 (complex64) / 2048 (complex128) amplitudes
  const int cap_lg = dtype == 0 ? 12 : 11;
  int lgS = std::max(0, cap_lg - g.k);
  int lg_rpw = 0;
  while (((int64_t)1 << lg_rpw) < g.rpw) ++lg_rpw;
  g.lgS = std::min(lgS, lg_rpw);
  g.mat = gw[G_MAT];
  g.slot = gw[G_SLOT];
  g.slot_base = slot_base;
  g.ntiles = ntiles;
}

hipError_t launch_wide(int dir, int dtype, const SvArgs& a, const WideGeom& g, int batch, hipStream_t st) {
  const dim3 grid((unsigned)g.ntiles, (unsigned)batch);
  const size_t sm = smem_bytes(dir, dtype, g);
  const void* fn = dir == 0 ? (dtype == 0 ? (const void*)wide_fwd_kernel<float> : (const void*)wide_fwd_kernel<double>)
                            : (dtype == 0 ? (const void*)wide_bwd_kernel<float> : (const void*)wide_bwd_kernel<double>);
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (dir == 0) {
    if (dtype == 0)
      hipLaunchKernelGGL(wide_fwd_kernel<float>, grid, dim3(NT), sm, st, a, g);
    else
      hipLaunchKernelGGL(wide_fwd_kernel<double>, grid, dim3(NT), sm, st, a, g);
  } else {
    if (dtype == 0)
      hipLaunchKernelGGL(wide_bwd_kernel<float>, grid, dim3(NT), sm, st, a, g);
    else
      hipLaunchKernelGGL(wide_bwd_kernel<double>, grid, dim3(NT), sm, st, a, g);
  }
  return hipGetLastError();
}

}  // namespace jq

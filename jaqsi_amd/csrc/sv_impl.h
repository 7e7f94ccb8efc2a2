// Tile-pass kernels for statevector (and density-matrix-as-vector) simulation
// on gfx950.
//
// One workgroup owns one tile of 2^T amplitudes of one circuit's vector: the
// amplitudes live in registers, 2^RB per thread, under a per-substep layout
// (register bits / thread bits -> tile bits).  Gates act on register bits
// only; control wires and diagonal factors may sit anywhere (register,
// thread or outside the tile).  Between substeps the tile moves through LDS.
// A pass therefore costs exactly one HBM read and one HBM write of the
// vector (two of each for the backward sweep), whatever its gate count.
//
// The forward kernel mirrors _run (jaqsi/simulation.py:400-409); the
// backward kernel mirrors the reversed-plan loop of _adjoint_expval
// (simulation.py:616-652): psi <- U^dagger psi (or the inverse of a
// non-unitary superoperator), bar[o,i] += lambda[o] psi[i], lambda <- U^T lambda.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jaqsi_amd.h"
#include "program.h"
#include "sv_kernels.h"

namespace jq {

template <typename R>
struct C2 {
  R x, y;
};

__device__ __forceinline__ uint32_t lds_swz(uint32_t x) { return x ^ ((x >> 5) & 31u); }

template <typename R>
__device__ __forceinline__ void cmul_acc(R& ar, R& ai, R br, R bi, R cr, R ci) {
  ar = fma(br, cr, ar);
  ar = fma(-bi, ci, ar);
  ai = fma(br, ci, ai);
  ai = fma(bi, cr, ai);
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// per-thread tile state
// ---------------------------------------------------------------------------
template <typename R, int RB, int NT>
struct Tile {
  static constexpr int NR = 1 << RB;
  static constexpr int LNT = __builtin_ctz(NT);
  R re[NR], im[NR];
};

struct Layout {
  uint32_t toff;   // tile-local offset of this thread
  uint64_t gthr;   // global offset of this thread's tile-local bits
  uint32_t rpos[8];  // tile-local bit of each register bit
  uint64_t gpos[8];  // global weight of each register bit
};

template <int RB, int LNT>
__device__ __forceinline__ void decode_layout(const int* sub, const int* tbits, Layout& L) {
  uint32_t toff = 0;
  uint64_t gthr = 0;
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < LNT; ++j) {
    const int tb = sub[1 + RB + j];
    if ((tid >> j) & 1) {
      toff |= 1u << tb;
      gthr |= 1ull << tbits[tb];
    }
  }
  L.toff = toff;
  L.gthr = gthr;
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    const int tb = sub[1 + b];
    L.rpos[b] = tb;
    L.gpos[b] = 1ull << tbits[tb];
  }
}

template <int RB>
__device__ __forceinline__ uint32_t reg_toff(const Layout& L, int i) {
  uint32_t x = L.toff;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((i >> b) & 1) x |= 1u << L.rpos[b];
  return x;
}

template <int RB>
__device__ __forceinline__ uint64_t reg_gidx(const Layout& L, uint64_t base, int i) {
  uint64_t g = base | L.gthr;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((i >> b) & 1) g |= L.gpos[b];
  return g;
}

// ---------------------------------------------------------------------------
// register-level gate application
// ---------------------------------------------------------------------------
// a 2x2 (or 4x4, 8x8) complex matrix held in registers
template <typename R, int D>
struct Mat {
  R r[D * D], i[D * D];
};

// Load a D x D matrix as M (mode 0), M^dagger (mode 1) or M^T (mode 2).
template <typename R, int D>
__device__ __forceinline__ void load_mat(const C2<R>* p, Mat<R, D>& m, int mode) {
  if (mode == 0) {
#pragma unroll
    for (int e = 0; e < D * D; ++e) {
      m.r[e] = p[e].x;
      m.i[e] = p[e].y;
    }
  } else {
    const R sg = mode == 1 ? R(-1) : R(1);
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int c = 0; c < D; ++c) {
        m.r[r * D + c] = p[c * D + r].x;
        m.i[r * D + c] = sg * p[c * D + r].y;
      }
  }
}

// Bits of the register index addressed by a gate, as template parameters.
// MC1: target A.  MC2: first target A (matrix MSB), second target B.
template <typename R, int NR, int A>
__device__ __forceinline__ void apply_mc1(R* re, R* im, const Mat<R, 2>& m, uint32_t creg, bool ok) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    if (i & (1 << A)) continue;
    const int j = i | (1 << A);
    const bool act = ok && ((uint32_t)(i & creg) == creg);
    const R x0r = re[i], x0i = im[i], x1r = re[j], x1i = im[j];
    R y0r = 0, y0i = 0, y1r = 0, y1i = 0;
    cmul_acc(y0r, y0i, m.r[0], m.i[0], x0r, x0i);
    cmul_acc(y0r, y0i, m.r[1], m.i[1], x1r, x1i);
    cmul_acc(y1r, y1i, m.r[2], m.i[2], x0r, x0i);
    cmul_acc(y1r, y1i, m.r[3], m.i[3], x1r, x1i);
    re[i] = act ? y0r : x0r;
    im[i] = act ? y0i : x0i;
    re[j] = act ? y1r : x1r;
    im[j] = act ? y1i : x1i;
  }
}

template <typename R, int NR, int A, int B>
__device__ __forceinline__ void apply_mc2(R* re, R* im, const Mat<R, 4>& m, uint32_t creg, bool ok) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    if (i & ((1 << A) | (1 << B))) continue;
    const bool act = ok && ((uint32_t)(i & creg) == creg);
    int idx[4] = {i, i | (1 << B), i | (1 << A), i | (1 << A) | (1 << B)};
    R xr[4], xi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      xr[q] = re[idx[q]];
      xi[q] = im[idx[q]];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      R ar = 0, ai = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) cmul_acc(ar, ai, m.r[r * 4 + c], m.i[r * 4 + c], xr[c], xi[c]);
      re[idx[r]] = act ? ar : xr[r];
      im[idx[r]] = act ? ai : xi[r];
    }
  }
}

template <typename R, int NR, int A>
__device__ __forceinline__ void apply_mcx(R* re, R* im, uint32_t creg, bool ok) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    if (i & (1 << A)) continue;
    const int j = i | (1 << A);
    const bool act = ok && ((uint32_t)(i & creg) == creg);
    const R ar = re[i], ai = im[i], br = re[j], bi = im[j];
    re[i] = act ? br : ar;
    im[i] = act ? bi : ai;
    re[j] = act ? ar : br;
    im[j] = act ? ai : bi;
  }
}

template <typename R, int NR, int A, int B>
__device__ __forceinline__ void apply_mcswap(R* re, R* im, uint32_t creg, bool ok) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    if (i & ((1 << A) | (1 << B))) continue;
    const int p = i | (1 << A), q = i | (1 << B);
    const bool act = ok && ((uint32_t)(i & creg) == creg);
    const R ar = re[p], ai = im[p], br = re[q], bi = im[q];
    re[p] = act ? br : ar;
    im[p] = act ? bi : ai;
    re[q] = act ? ar : br;
    im[q] = act ? ai : bi;
  }
}

// Dense 3-wire block on register bits 2..0 (wire j <-> register bit 2-j),
// evaluated through LDS so the 8x8 matrix never occupies registers.
template <typename R, int RB>
__device__ __forceinline__ void apply_dense3(R* re, R* im, const C2<R>* mp, int mode, const Layout& L, C2<R>* lds) {
  constexpr int NR = 1 << RB;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    C2<R> v;
    v.x = re[i];
    v.y = im[i];
    lds[lds_swz(reg_toff<RB>(L, i))] = v;
  }
  __syncthreads();
  const R sg = mode == 1 ? R(-1) : R(1);
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const int r = i & 7;
    R ar = 0, ai = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const C2<R> x = lds[lds_swz(reg_toff<RB>(L, (i & ~7) | c))];
      const C2<R> mv = mode == 0 ? mp[r * 8 + c] : mp[c * 8 + r];
      cmul_acc(ar, ai, mv.x, sg * mv.y, x.x, x.y);
    }
    re[i] = ar;
    im[i] = ai;
  }
  __syncthreads();
}

// Diagonal index of register i: uniform part + register bits.
struct DiagLoc {
  uint32_t uni;     // contribution of thread / outside bits
  int rbit[5];      // register bit of wire j, -1 if not a register bit
  int k;
};

__device__ __forceinline__ void decode_diag(const int* g, const Layout& L, uint64_t base, DiagLoc& dl) {
  const int k = g[G_K];
  dl.k = k;
  dl.uni = 0;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    dl.rbit[j] = -1;
    if (j >= k) continue;
    const int loc = (g[G_DW0 + j / 4] >> (8 * (j % 4))) & 0xff;
    const int type = loc >> 6, pos = loc & 63;
    uint32_t bit = 0;
    if (type == DL_REG)
      dl.rbit[j] = pos;
    else if (type == DL_THR)
      bit = (L.toff >> pos) & 1u;
    else
      bit = (uint32_t)((base >> pos) & 1ull);
    dl.uni |= bit << (k - 1 - j);
  }
}

__device__ __forceinline__ uint32_t diag_index(const DiagLoc& dl, int i) {
  uint32_t x = dl.uni;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    if (j < dl.k && dl.rbit[j] >= 0) x |= (uint32_t)((i >> dl.rbit[j]) & 1) << (dl.k - 1 - j);
  return x;
}

// mode 0: d, 1: conj(d) (U^dagger), 2: d (U^T)
template <typename R, int NR>
__device__ __forceinline__ void apply_diag(R* re, R* im, const C2<R>* d, const DiagLoc& dl, int mode) {
  if (dl.k <= 2) {
    R dr[4], di[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const C2<R> v = d[q < (1 << dl.k) ? q : 0];
      dr[q] = v.x;
      di[q] = mode == 1 ? -v.y : v.y;
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const uint32_t x = diag_index(dl, i);
      R fr = dr[0], fi = di[0];
      fr = x == 1 ? dr[1] : fr;
      fi = x == 1 ? di[1] : fi;
      fr = x == 2 ? dr[2] : fr;
      fi = x == 2 ? di[2] : fi;
      fr = x == 3 ? dr[3] : fr;
      fi = x == 3 ? di[3] : fi;
      const R ar = re[i], ai = im[i];
      re[i] = ar * fr - ai * fi;
      im[i] = ar * fi + ai * fr;
    }
  } else {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const uint32_t x = diag_index(dl, i);
      const C2<R> v = d[x];
      const R fr = v.x, fi = mode == 1 ? -v.y : v.y;
      const R ar = re[i], ai = im[i];
      re[i] = ar * fr - ai * fi;
      im[i] = ar * fi + ai * fr;
    }
  }
}

// ---------------------------------------------------------------------------
// cotangent accumulation: bar[o][i] += lambda[o] psi[i] (no conjugation)
// ---------------------------------------------------------------------------
template <typename R, int NR, int A>
__device__ __forceinline__ void grad_mc1(const R* pr, const R* pi, const R* lr, const R* li, uint32_t creg, bool ok,
                                          R* acc) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    if (i & (1 << A)) continue;
    const int j = i | (1 << A);
    const bool act = ok && ((uint32_t)(i & creg) == creg);
    const int id[2] = {i, j};
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        R ar = 0, ai = 0;
        cmul_acc(ar, ai, lr[id[o]], li[id[o]], pr[id[c]], pi[id[c]]);
        acc[2 * (o * 2 + c)] += act ? ar : R(0);
        acc[2 * (o * 2 + c) + 1] += act ? ai : R(0);
      }
  }
}

template <typename R>
__device__ __forceinline__ void flush_vals(const R* acc, int nv, double* slots, int base);

template <typename R, int NR, int A, int B>
__device__ __forceinline__ void grad_mc2(const R* pr, const R* pi, const R* lr, const R* li, uint32_t creg, bool ok,
                                          double* slots, int slot) {
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    R acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = 0;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      if (i & ((1 << A) | (1 << B))) continue;
      const bool act = ok && ((uint32_t)(i & creg) == creg);
      const int id[4] = {i, i | (1 << B), i | (1 << A), i | (1 << A) | (1 << B)};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        R ar = 0, ai = 0;
        cmul_acc(ar, ai, lr[id[o]], li[id[o]], pr[id[c]], pi[id[c]]);
        acc[2 * c] += act ? ar : R(0);
        acc[2 * c + 1] += act ? ai : R(0);
      }
    }
    flush_vals<R>(acc, 8, slots, slot + 8 * o);
  }
}

template <typename R, int NR>
__device__ __forceinline__ void grad_diag(const R* pr, const R* pi, const R* lr, const R* li, const DiagLoc& dl,
                                           R* acc) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const uint32_t x = diag_index(dl, i);
    R ar = 0, ai = 0;
    cmul_acc(ar, ai, lr[i], li[i], pr[i], pi[i]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc[2 * q] += x == (uint32_t)q ? ar : R(0);
      acc[2 * q + 1] += x == (uint32_t)q ? ai : R(0);
    }
  }
}

// Reduce per-thread accumulators over the workgroup into LDS slots (double).
template <typename R>
__device__ __forceinline__ void flush_vals(const R* acc, int nv, double* slots, int base) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    if (e >= nv) break;
    const R v = wave_sum<R>(acc[e]);
    if ((threadIdx.x & 63) == 0) atomicAdd(&slots[base + e], (double)v);
  }
}

// ---------------------------------------------------------------------------
// dispatch on register bit positions
// ---------------------------------------------------------------------------
#define JQ_CASE1(F, ...)                      \
  switch (ra) {                               \
    case 0: F<R, NR, 0>(__VA_ARGS__); break;  \
    case 1: F<R, NR, 1>(__VA_ARGS__); break;  \
    case 2: F<R, NR, 2>(__VA_ARGS__); break;  \
    case 3: F<R, NR, 3>(__VA_ARGS__); break;  \
    default: if constexpr (RB > 4) F<R, NR, (RB > 4 ? 4 : 0)>(__VA_ARGS__); break; \
  }

#define JQ_PAIR(F, A_, B_, ...) \
  case (A_ * 8 + B_): if constexpr (A_ < RB && B_ < RB) F<R, NR, A_, B_>(__VA_ARGS__); break;

#define JQ_CASE2(F, ...)                                                                          \
  switch (ra * 8 + rb) {                                                                          \
    JQ_PAIR(F, 0, 1, __VA_ARGS__) JQ_PAIR(F, 0, 2, __VA_ARGS__) JQ_PAIR(F, 0, 3, __VA_ARGS__)     \
    JQ_PAIR(F, 0, 4, __VA_ARGS__) JQ_PAIR(F, 1, 0, __VA_ARGS__) JQ_PAIR(F, 1, 2, __VA_ARGS__)     \
    JQ_PAIR(F, 1, 3, __VA_ARGS__) JQ_PAIR(F, 1, 4, __VA_ARGS__) JQ_PAIR(F, 2, 0, __VA_ARGS__)     \
    JQ_PAIR(F, 2, 1, __VA_ARGS__) JQ_PAIR(F, 2, 3, __VA_ARGS__) JQ_PAIR(F, 2, 4, __VA_ARGS__)     \
    JQ_PAIR(F, 3, 0, __VA_ARGS__) JQ_PAIR(F, 3, 1, __VA_ARGS__) JQ_PAIR(F, 3, 2, __VA_ARGS__)     \
    JQ_PAIR(F, 3, 4, __VA_ARGS__) JQ_PAIR(F, 4, 0, __VA_ARGS__) JQ_PAIR(F, 4, 1, __VA_ARGS__)     \
    JQ_PAIR(F, 4, 2, __VA_ARGS__) JQ_PAIR(F, 4, 3, __VA_ARGS__)                                   \
    default: break;                                                                               \
  }

struct GateCtx {
  uint32_t creg;
  bool ok;  // thread / outside control bits satisfied
};

__device__ __forceinline__ GateCtx gate_ctx(const int* g, const Layout& L, uint64_t base) {
  GateCtx c;
  c.creg = (uint32_t)g[G_CREG];
  const uint32_t cthr = (uint32_t)g[G_CTHR];
  const uint64_t cout = ((uint64_t)(uint32_t)g[G_COUT_HI] << 32) | (uint32_t)g[G_COUT_LO];
  c.ok = ((L.toff & cthr) == cthr) && ((base & cout) == cout);
  return c;
}

// Apply gate record g to one register state.  mode 0: U, 1: U^dagger (or the
// stored inverse), 2: U^T.
template <typename R, int RB>
__device__ __forceinline__ void apply_gate(R* re, R* im, const int* g, const C2<R>* pool, const Layout& L, uint64_t base, int mode,
                           C2<R>* lds) {
  constexpr int NR = 1 << RB;
  const int kind = g[G_KIND];
  const int ra = g[G_RA], rb = g[G_RB];
  const GateCtx cx = gate_ctx(g, L, base);
  const bool use_inv = (mode == 1) && g[G_INV] >= 0;
  const C2<R>* mp = pool + (use_inv ? g[G_INV] : g[G_MAT]);
  const int mm = use_inv ? 0 : mode;
  switch (kind) {
    case AK_MC1: {
      Mat<R, 2> m;
      load_mat<R, 2>(mp, m, mm);
      JQ_CASE1(apply_mc1, re, im, m, cx.creg, cx.ok)
      break;
    }
    case AK_MC2: {
      Mat<R, 4> m;
      load_mat<R, 4>(mp, m, mm);
      JQ_CASE2(apply_mc2, re, im, m, cx.creg, cx.ok)
      break;
    }
    case AK_MCX: {
      JQ_CASE1(apply_mcx, re, im, cx.creg, cx.ok)
      break;
    }
    case AK_MCSWAP: {
      JQ_CASE2(apply_mcswap, re, im, cx.creg, cx.ok)
      break;
    }
    case AK_DIAG: {
      DiagLoc dl;
      decode_diag(g, L, base, dl);
      apply_diag<R, NR>(re, im, mp, dl, use_inv ? 0 : mode);
      break;
    }
    case AK_DENSEK: {
      if constexpr (RB >= 3) apply_dense3<R, RB>(re, im, mp, mm, L, lds);
      break;
    }
    default:
      break;
  }
}

template <typename R, int RB>
__device__ __forceinline__ void accumulate_grad(const R* pr, const R* pi, const R* lr, const R* li, const int* g, const Layout& L,
                                uint64_t base, double* slots) {
  constexpr int NR = 1 << RB;
  const int kind = g[G_KIND];
  const int ra = g[G_RA], rb = g[G_RB];
  const int slot = g[G_SLOT];
  const GateCtx cx = gate_ctx(g, L, base);
  if (kind == AK_MC1) {
    R acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = 0;
    JQ_CASE1(grad_mc1, pr, pi, lr, li, cx.creg, cx.ok, acc)
    flush_vals<R>(acc, 8, slots, slot);
  } else if (kind == AK_MC2) {
    JQ_CASE2(grad_mc2, pr, pi, lr, li, cx.creg, cx.ok, slots, slot)
  } else if (kind == AK_DIAG) {
    DiagLoc dl;
    decode_diag(g, L, base, dl);
    R acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = 0;
    grad_diag<R, NR>(pr, pi, lr, li, dl, acc);
    flush_vals<R>(acc, 2 << dl.k, slots, slot);
  }
}

// ---------------------------------------------------------------------------
// global <-> register, register <-> LDS
// ---------------------------------------------------------------------------
template <typename R, int RB>
__device__ __forceinline__ void load_global(R* re, R* im, const C2<R>* src, uint64_t vec_len, const Layout& L,
                                            uint64_t base, bool zero_init) {
  constexpr int NR = 1 << RB;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const uint64_t g = reg_gidx<RB>(L, base, i);
    if (zero_init) {
      re[i] = g == 0 ? R(1) : R(0);
      im[i] = 0;
    } else if (g < vec_len) {
      const C2<R> v = src[g];
      re[i] = v.x;
      im[i] = v.y;
    } else {
      re[i] = 0;
      im[i] = 0;
    }
  }
}

template <typename R, int RB>
__device__ __forceinline__ void store_global(const R* re, const R* im, C2<R>* dst, uint64_t vec_len, const Layout& L,
                                             uint64_t base) {
  constexpr int NR = 1 << RB;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const uint64_t g = reg_gidx<RB>(L, base, i);
    if (g < vec_len) {
      C2<R> v;
      v.x = re[i];
      v.y = im[i];
      dst[g] = v;
    }
  }
}

template <typename R, int RB>
__device__ __forceinline__ void relayout(R* re, R* im, C2<R>* lds, const Layout& from, const Layout& to) {
  constexpr int NR = 1 << RB;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    C2<R> v;
    v.x = re[i];
    v.y = im[i];
    lds[lds_swz(reg_toff<RB>(from, i))] = v;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const C2<R> v = lds[lds_swz(reg_toff<RB>(to, i))];
    re[i] = v.x;
    im[i] = v.y;
  }
  __syncthreads();
}

__device__ __forceinline__ uint64_t tile_base(int tile, const int* tbits, int T, int n) {
  // deposit the tile number into the real global bits that are not tile bits
  uint64_t mask = 0;
  for (int j = 0; j < T; ++j)
    if (tbits[j] < n) mask |= 1ull << tbits[j];
  uint64_t base = 0;
  int src = 0;
  for (int gb = 0; gb < n; ++gb) {
    if (mask & (1ull << gb)) continue;
    if ((tile >> src) & 1) base |= 1ull << gb;
    ++src;
  }
  return base;
}

// ---------------------------------------------------------------------------
// observables
// ---------------------------------------------------------------------------
// weight of amplitude g for the lambda seed: sum_o ct_o d_o[bit]
__device__ __forceinline__ double obs_weight(uint64_t g, const SvArgs& a, const double* ct) {
  double w = 0;
  for (int o = 0; o < a.n_obs; ++o) {
    const int bit = (int)((g >> a.obs_gbit[o]) & 1ull);
    w += ct[o] * a.obs_d[2 * o + bit];
  }
  return w;
}

template <typename R, int RB>
__device__ __forceinline__ void expval_partials(const R* re, const R* im, const SvArgs& a, const Layout& L, uint64_t base,
                                double* ev_lds) {
  constexpr int NR = 1 << RB;
  const int nq = a.n_wires / 2;
  const uint64_t lowmask = (1ull << nq) - 1;
  for (int o = 0; o < a.n_obs; ++o) {
    const int gb = a.obs_gbit[o];
    const double d0 = a.obs_d[2 * o], d1 = a.obs_d[2 * o + 1];
    R acc = 0;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const uint64_t g = reg_gidx<RB>(L, base, i);
      const R d = ((g >> gb) & 1ull) ? (R)d1 : (R)d0;
      R v;
      if (a.obs_kind == JQ_OBS_SV_DIAG)
        v = re[i] * re[i] + im[i] * im[i];
      else
        v = ((g >> nq) == (g & lowmask)) ? re[i] : R(0);
      acc = fma(v, d, acc);
    }
    const R s = wave_sum<R>(acc);
    if ((threadIdx.x & 63) == 0) atomicAdd(&ev_lds[o], (double)s);
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <typename R, int RB, int NT>
__device__ __forceinline__ void forward_tile(const SvArgs& a, int pass_idx, int circ, int tile, R* re, R* im, Layout& L,
                             uint64_t& base, C2<R>* lds, bool do_load) {
  const int* P = a.prog;
  const int T = P[2];
  const int* pass = P + P[HDR_WORDS + pass_idx];
  const int nsub = pass[0];
  const int flags = pass[1];
  const int* tbits = pass + 4;
  constexpr int LNT = __builtin_ctz(NT);
  base = tile_base(tile, tbits, T, a.n_wires);
  const C2<R>* pool = reinterpret_cast<const C2<R>*>(a.mats) + (int64_t)circ * a.pool_stride;
  const int* soff = pass + 4 + T;
  decode_layout<RB, LNT>(P + soff[0], tbits, L);
  (void)flags;
  if (do_load) {
    const bool zero = a.src == nullptr;
    const C2<R>* src = reinterpret_cast<const C2<R>*>(a.src);
    load_global<R, RB>(re, im, zero ? nullptr : src + (int64_t)circ * a.src_stride, a.vec_len, L, base, zero);
  }
  for (int s = 0; s < nsub; ++s) {
    const int* sub = P + soff[s];
    const int ng = sub[0];
    if (s > 0) {
      Layout nl;
      decode_layout<RB, LNT>(sub, tbits, nl);
      relayout<R, RB>(re, im, lds, L, nl);
      L = nl;
    }
    const int* g = sub + 1 + T;
    for (int q = 0; q < ng; ++q, g += GW) apply_gate<R, RB>(re, im, g, pool, L, base, 0, lds);
  }
}

template <typename R, int RB, int NT>
__global__ __launch_bounds__(NT) void sv_forward_kernel(SvArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int T = a.prog[2];
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);
  double* ev = reinterpret_cast<double*>(smem + (sizeof(C2<R>) << T));
  const int circ = blockIdx.y;
  const int tile = blockIdx.x;
  const int* pass = a.prog + a.prog[HDR_WORDS + a.pass];
  const bool last = (pass[1] & PF_LAST) != 0;
  const bool want_ev = last && a.obs_kind != JQ_OBS_NONE && a.ev_part != nullptr;
  if (want_ev)
    for (int o = threadIdx.x; o < a.n_obs; o += NT) ev[o] = 0;
  R re[1 << RB], im[1 << RB];
  Layout L;
  uint64_t base;
  forward_tile<R, RB, NT>(a, a.pass, circ, tile, re, im, L, base, lds, true);
  store_global<R, RB>(re, im, reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride, a.vec_len, L, base);
  if (want_ev) {
    __syncthreads();
    expval_partials<R, RB>(re, im, a, L, base, ev);
    __syncthreads();
    for (int o = threadIdx.x; o < a.n_obs; o += NT)
      a.ev_part[((int64_t)circ * gridDim.x + tile) * a.n_obs + o] = ev[o];
  }
}

// Backward sweep over one pass (reverse substeps, reverse gates).
// psi: loaded from a.src (the pass output), lambda from a.lam or seeded.
template <typename R, int RB, int NT>
__device__ __forceinline__ void backward_tile(const SvArgs& a, int pass_idx, int circ, int tile, R* pr, R* pi, R* lr, R* li,
                              Layout& L, uint64_t& base, C2<R>* lds, double* slots, bool loaded) {
  const int* P = a.prog;
  const int T = P[2];
  const int* pass = P + P[HDR_WORDS + pass_idx];
  const int nsub = pass[0];
  const int flags = pass[1];
  const int* tbits = pass + 4;
  constexpr int LNT = __builtin_ctz(NT);
  constexpr int NR = 1 << RB;
  base = tile_base(tile, tbits, T, a.n_wires);
  const C2<R>* pool = reinterpret_cast<const C2<R>*>(a.mats) + (int64_t)circ * a.pool_stride;
  const int* soff = pass + 4 + T;
  decode_layout<RB, LNT>(P + soff[nsub - 1], tbits, L);
  if (!loaded) {
    const C2<R>* ps = reinterpret_cast<const C2<R>*>(a.src) + (int64_t)circ * a.src_stride;
    load_global<R, RB>(pr, pi, ps, a.vec_len, L, base, false);
    if (flags & PF_LAST) {
      // lambda seed: JAX cotangent of the measurement (simulation.py:617)
      const double* ct = a.ct + (int64_t)circ * a.n_obs;
      const int nq = a.n_wires / 2;
      const uint64_t lowmask = (1ull << nq) - 1;
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        const uint64_t g = reg_gidx<RB>(L, base, i);
        const double w = obs_weight(g, a, ct);
        if (a.obs_kind == JQ_OBS_SV_DIAG) {
          lr[i] = (R)(2.0 * w) * pr[i];
          li[i] = -(R)(2.0 * w) * pi[i];
        } else {
          lr[i] = ((g >> nq) == (g & lowmask) && g < a.vec_len) ? (R)w : R(0);
          li[i] = 0;
        }
      }
    } else {
      const C2<R>* ls = reinterpret_cast<const C2<R>*>(a.lam) + (int64_t)circ * a.lam_stride;
      load_global<R, RB>(lr, li, ls, a.vec_len, L, base, false);
    }
  }
  for (int s = nsub - 1; s >= 0; --s) {
    const int* sub = P + soff[s];
    if (s < nsub - 1) {
      Layout nl;
      decode_layout<RB, LNT>(sub, tbits, nl);
      relayout<R, RB>(pr, pi, lds, L, nl);
      relayout<R, RB>(lr, li, lds, L, nl);
      L = nl;
    }
    const int ng = sub[0];
    const int* g0 = sub + 1 + T;
    for (int q = ng - 1; q >= 0; --q) {
      const int* g = g0 + q * GW;
      apply_gate<R, RB>(pr, pi, g, pool, L, base, 1, lds);
      if (g[G_SLOT] >= 0 && slots != nullptr) accumulate_grad<R, RB>(pr, pi, lr, li, g, L, base, slots);
      apply_gate<R, RB>(lr, li, g, pool, L, base, 2, lds);
    }
  }
}

template <typename R, int RB, int NT>
__global__ __launch_bounds__(NT) void sv_backward_kernel(SvArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int T = a.prog[2];
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);
  double* slots = reinterpret_cast<double*>(smem + (sizeof(C2<R>) << T));
  const int circ = blockIdx.y;
  const int tile = blockIdx.x;
  const int* pass = a.prog + a.prog[HDR_WORDS + a.pass];
  const int nslots = pass[2];
  const int slot_base = pass[3];
  for (int e = threadIdx.x; e < nslots; e += NT) slots[e] = 0;
  __syncthreads();
  R pr[1 << RB], pi[1 << RB], lr[1 << RB], li[1 << RB];
  Layout L;
  uint64_t base;
  backward_tile<R, RB, NT>(a, a.pass, circ, tile, pr, pi, lr, li, L, base, lds, slots, false);
  C2<R>* lam = reinterpret_cast<C2<R>*>(a.lam) + (int64_t)circ * a.lam_stride;
  store_global<R, RB>(lr, li, lam, a.vec_len, L, base);
  if (a.dst != nullptr)
    store_global<R, RB>(pr, pi, reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride, a.vec_len, L, base);
  __syncthreads();
  double* gp = a.grad_part + (int64_t)circ * a.part_stride + slot_base + (int64_t)tile * nslots;
  for (int e = threadIdx.x; e < nslots; e += NT) gp[e] = slots[e];
}

// Whole circuit in one workgroup (the vector fits one tile): forward, the
// expectation values, the lambda seed and the backward sweep without the
// vector ever leaving registers.
template <typename R, int RB, int NT>
__global__ __launch_bounds__(NT) void sv_fused_kernel(SvArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int T = a.prog[2];
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);
  double* ev = reinterpret_cast<double*>(smem + (sizeof(C2<R>) << T));
  const int* pass = a.prog + a.prog[HDR_WORDS + 0];
  const int nslots = pass[2];
  double* slots = ev + 64;
  const int circ = blockIdx.y;
  for (int e = threadIdx.x; e < nslots; e += NT) slots[e] = 0;
  for (int o = threadIdx.x; o < 64; o += NT) ev[o] = 0;
  __syncthreads();
  constexpr int NR = 1 << RB;
  R pr[NR], pi[NR], lr[NR], li[NR];
  Layout L;
  uint64_t base;
  forward_tile<R, RB, NT>(a, 0, circ, 0, pr, pi, L, base, lds, true);
  if (a.dst != nullptr)
    store_global<R, RB>(pr, pi, reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride, a.vec_len, L, base);
  if (a.ev_part != nullptr) {
    expval_partials<R, RB>(pr, pi, a, L, base, ev);
  }
  // seed lambda under the last forward layout
  {
    const double* ct = a.ct + (int64_t)circ * a.n_obs;
    const int nq = a.n_wires / 2;
    const uint64_t lowmask = (1ull << nq) - 1;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const uint64_t g = reg_gidx<RB>(L, base, i);
      const double w = obs_weight(g, a, ct);
      if (a.obs_kind == JQ_OBS_SV_DIAG) {
        lr[i] = (R)(2.0 * w) * pr[i];
        li[i] = -(R)(2.0 * w) * pi[i];
      } else {
        lr[i] = ((g >> nq) == (g & lowmask) && g < a.vec_len) ? (R)w : R(0);
        li[i] = 0;
      }
    }
  }
  backward_tile<R, RB, NT>(a, 0, circ, 0, pr, pi, lr, li, L, base, lds, slots, true);
  if (a.lam != nullptr)
    store_global<R, RB>(lr, li, reinterpret_cast<C2<R>*>(a.lam) + (int64_t)circ * a.lam_stride, a.vec_len, L, base);
  __syncthreads();
  double* gp = a.grad_part + (int64_t)circ * a.part_stride + pass[3];
  for (int e = threadIdx.x; e < nslots; e += NT) gp[e] = slots[e];
  if (a.ev_part != nullptr)
    for (int o = threadIdx.x; o < a.n_obs; o += NT) a.ev_part[(int64_t)circ * a.n_obs + o] = ev[o];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename R, int RB, int NT>
hipError_t launch_cfg(int which, const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  dim3 grid(ntiles, batch);
  if (which == 0) {
    (void)hipFuncSetAttribute((const void*)sv_forward_kernel<R, RB, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipLaunchKernelGGL((sv_forward_kernel<R, RB, NT>), grid, dim3(NT), smem, st, a);
  } else if (which == 1) {
    (void)hipFuncSetAttribute((const void*)sv_backward_kernel<R, RB, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipLaunchKernelGGL((sv_backward_kernel<R, RB, NT>), grid, dim3(NT), smem, st, a);
  } else {
    (void)hipFuncSetAttribute((const void*)sv_fused_kernel<R, RB, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipLaunchKernelGGL((sv_fused_kernel<R, RB, NT>), grid, dim3(NT), smem, st, a);
  }
  return hipGetLastError();
}

}  // namespace jq

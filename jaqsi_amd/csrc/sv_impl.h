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
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "jaqsi_amd.h"
#include "program.h"
#include "sv_kernels.h"

namespace jq {

template <typename R>
struct C2 {
  R x, y;
};

// Tile bits of the pass kernels: 2^T amplitudes per tile (64 KiB).
template <typename R>
constexpr int kTileBits = sizeof(R) == 4 ? 13 : 12;

// LDS state regions of a backward workgroup on a tile of TB bits with 2^RB
// amplitudes per thread: two (psi and lambda side by side) while both fit in
// 128 KiB and the workgroup has 512 or more threads, otherwise one, through
// which psi and lambda are exchanged in turn: 2^14 complex64 amplitudes fill
// 128 KiB, and workgroups of fewer threads (32-amplitude register tiles on
// 13-bit tiles) share the CU with a second workgroup.
template <typename R, int TB, int RB>
constexpr int kRegions = (2 * (sizeof(C2<R>) << TB) <= 128 * 1024 && TB - RB >= 9) ? 2 : 1;
template <typename R, int RB>
__device__ __forceinline__ int lds_regions(int T) {
  return (2 * (sizeof(C2<R>) << T) <= 128 * 1024 && T - RB >= 9) ? 2 : 1;
}

// Program words, block matrices and observable tables are read-only for the
// whole launch: reading them through the constant address space lets every
// wave-uniform access compile to a scalar load (s_load into SGPRs) instead of
// a vector load whose result occupies VGPRs.
#if defined(__HIP_DEVICE_COMPILE__)
#define JQ_CONST __attribute__((address_space(4)))
#define JQ_LDS __attribute__((address_space(3)))
#else
#define JQ_CONST
#define JQ_LDS
#endif
using cint = const JQ_CONST int;
template <typename R>
using cmat = const JQ_CONST C2<R>;

__device__ __forceinline__ uint32_t lds_swz(uint32_t x) { return x ^ ((x >> 5) & 31u); }

template <typename R>
__device__ __forceinline__ void cmul_acc(R& ar, R& ai, R br, R bi, R cr, R ci) {
  ar = fma(br, cr, ar);
  ar = fma(-bi, ci, ar);
  ai = fma(br, ci, ai);
  ai = fma(bi, cr, ai);
}

// Sum over the 64 lanes of a wave, returned wave-uniform.  Within each row
// of 16 lanes the DPP network (quad_perm xor 1, xor 2, half-row mirror, row
// mirror) leaves every lane with its row sum; the four row sums are then
// combined with readlane.  All lanes of the wave must be active.
__device__ __forceinline__ float dpp_row_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));  // quad_perm(1,0,3,2)
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));  // quad_perm(2,3,0,1)
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false)); // row_half_mirror
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false)); // row_mirror
  return v;
}

// x + (x moved by DPP control CTRL): one step of a DPP reduction.
template <int CTRL>
__device__ __forceinline__ double dpp_step_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return x + __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double dpp_row_sum(double v) {
  // move both 32-bit halves through the same network, add as doubles
  v = dpp_step_d<0xB1>(v);
  v = dpp_step_d<0x4E>(v);
  v = dpp_step_d<0x141>(v);
  v = dpp_step_d<0x140>(v);
  return v;
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
  v = dpp_row_sum(v);
  if constexpr (sizeof(R) == 4) {
    const float a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (a + b) + (c + d);
  } else {
    auto rl = [](double x, int lane) {
      const long long bb = __double_as_longlong(x);
      const int lo = __builtin_amdgcn_readlane((int)(bb & 0xffffffffll), lane);
      const int hi = __builtin_amdgcn_readlane((int)(bb >> 32), lane);
      return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    };
    return (rl(v, 0) + rl(v, 16)) + (rl(v, 32) + rl(v, 48));
  }
}

// Add a wave-uniform value into an LDS double from one lane.
// Add a wave-uniform value once per wave (lane 0 issues the atomic).
__device__ __forceinline__ void lds_add(double* p, double v) {
  if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Per-lane LDS atomic add (each calling lane its own address and value).
__device__ __forceinline__ void lds_add_lane(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
  uint32_t rpos[8];  // tile-local bit of each register bit (wave-uniform)
  cint* tbits;       // global bit of each tile bit (the pass record)
};

// Global weight of register bit b (wave-uniform; read from the pass record
// instead of being kept live, which keeps the layouts small in SGPRs).
__device__ __forceinline__ uint64_t gpos(const Layout& L, int b) { return 1ull << L.tbits[L.rpos[b]]; }

// Substep layouts only need the tile-local thread offset: the global offset
// gthr is read exclusively under the canonical layout (loads, stores,
// expectation values, lambda seed), which canonical_layout sets up.
template <int RB, int LNT, typename SubPtr>
__device__ __forceinline__ void decode_layout(SubPtr sub, cint* tbits, Layout& L) {
  uint32_t toff = 0;
  const uint32_t tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < LNT; ++j) toff |= ((tid >> j) & 1u) << sub[1 + RB + j];
  L.toff = toff;
  L.gthr = 0;
  L.tbits = tbits;
#pragma unroll
  for (int b = 0; b < RB; ++b) L.rpos[b] = sub[1 + b];
}

template <int RB>
__device__ __forceinline__ uint32_t reg_toff(const Layout& L, int i) {
  uint32_t x = L.toff;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((i >> b) & 1) x |= 1u << L.rpos[b];
  return x;
}

// Tile-local offset of register i alone (wave-uniform).
template <int RB>
__device__ __forceinline__ uint32_t reg_roff(const Layout& L, int i) {
  uint32_t x = 0;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((i >> b) & 1) x |= 1u << L.rpos[b];
  return x;
}

// Global offset of register i alone (wave-uniform).
template <int RB>
__device__ __forceinline__ uint64_t reg_goff(const Layout& L, int i) {
  uint64_t g = 0;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((i >> b) & 1) g |= gpos(L, b);
  return g;
}

template <int RB>
__device__ __forceinline__ uint64_t reg_gidx(const Layout& L, uint64_t base, int i) {
  uint64_t g = base | L.gthr;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((i >> b) & 1) g |= gpos(L, b);
  return g;
}

// ---------------------------------------------------------------------------
// register-level gate application
// ---------------------------------------------------------------------------
// A small complex matrix held in (scalar) registers.
template <typename R, int D>
struct Mat {
  R r[D * D], i[D * D];
};

template <typename R>
__device__ __forceinline__ void load2(cmat<R>* p, Mat<R, 2>& m) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    m.r[e] = p[e].x;
    m.i[e] = p[e].y;
  }
}

// 4x4 matrix; `swap` exchanges the roles of the two target bits, so the
// kernels only need the register-bit pairs A > B.
template <typename R>
__device__ __forceinline__ void load4(cmat<R>* p, Mat<R, 4>& m, bool swap) {
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int sr = swap ? (((r & 1) << 1) | (r >> 1)) : r;
      const int sc = swap ? (((c & 1) << 1) | (c >> 1)) : c;
      m.r[r * 4 + c] = p[sr * 4 + sc].x;
      m.i[r * 4 + c] = p[sr * 4 + sc].y;
    }
}

// conjugate transpose (the inverse of a unitary)
template <typename R, int D>
__device__ __forceinline__ void adjoint_of(const Mat<R, D>& a, Mat<R, D>& o) {
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) {
      o.r[r * D + c] = a.r[c * D + r];
      o.i[r * D + c] = -a.i[c * D + r];
    }
}

// y = M x for a D-vector held in registers at indices idx[]
template <typename R, int D>
__device__ __forceinline__ void mv(const Mat<R, D>& m, const R* xr, const R* xi, R* yr, R* yi) {
#pragma unroll
  for (int r = 0; r < D; ++r) {
    R ar = xr[0] * m.r[r * D];
    R ai = xi[0] * m.r[r * D];
    ar = fma(-m.i[r * D], xi[0], ar);
    ai = fma(m.i[r * D], xr[0], ai);
#pragma unroll
    for (int c = 1; c < D; ++c) cmul_acc(ar, ai, m.r[r * D + c], m.i[r * D + c], xr[c], xi[c]);
    yr[r] = ar;
    yi[r] = ai;
  }
}

// y = M^T x
template <typename R, int D>
__device__ __forceinline__ void mvt(const Mat<R, D>& m, const R* xr, const R* xi, R* yr, R* yi) {
#pragma unroll
  for (int c = 0; c < D; ++c) {
    R ar = xr[0] * m.r[c];
    R ai = xi[0] * m.r[c];
    ar = fma(-m.i[c], xi[0], ar);
    ai = fma(m.i[c], xr[0], ai);
#pragma unroll
    for (int r = 1; r < D; ++r) cmul_acc(ar, ai, m.r[r * D + c], m.i[r * D + c], xr[r], xi[r]);
    yr[c] = ar;
    yi[c] = ai;
  }
}

// Register index group of a gate: MC1 target A; MC2 targets A > B with the
// matrix index (bit A, bit B).  creg is a wave-uniform mask over register
// bits: groups whose register controls are not all one are skipped by a
// scalar branch.  Controls on thread bits / outside the tile are tested by
// the caller.
template <int NR, int A>
struct G1 {
  static constexpr int n = NR / 2;
  __device__ static constexpr int base(int k) { return ((k >> A) << (A + 1)) | (k & ((1 << A) - 1)); }
  __device__ static constexpr int at(int k, int q) { return base(k) | (q << A); }
};
template <int NR, int A, int B>
struct G2 {
  static constexpr int n = NR / 4;
  __device__ static constexpr int base(int k) {
    const int lo = ((k >> B) << (B + 1)) | (k & ((1 << B) - 1));
    return ((lo >> A) << (A + 1)) | (lo & ((1 << A) - 1));
  }
  __device__ static constexpr int at(int k, int q) { return base(k) | ((q >> 1) << A) | ((q & 1) << B); }
};

template <typename R, int NR, int D, typename G, bool CTRL>
__device__ __forceinline__ void fw_dense(R* re, R* im, const Mat<R, D>& m, uint32_t creg) {
#pragma unroll
  for (int k = 0; k < G::n; ++k) {
    if (CTRL && ((uint32_t)G::base(k) & creg) != creg) continue;
    R xr[D], xi[D], yr[D], yi[D];
#pragma unroll
    for (int q = 0; q < D; ++q) {
      xr[q] = re[G::at(k, q)];
      xi[q] = im[G::at(k, q)];
    }
    mv<R, D>(m, xr, xi, yr, yi);
#pragma unroll
    for (int q = 0; q < D; ++q) {
      re[G::at(k, q)] = yr[q];
      im[G::at(k, q)] = yi[q];
    }
  }
}

// Backward step of one dense block on one register group set:
//   psi <- Ui psi ; acc[o][c] += lambda[o] psi[c] ; lambda <- U^T lambda
template <typename R, int NR, int D, typename G, bool CTRL>
__device__ __forceinline__ void bw_dense(R* pr, R* pi, R* lr, R* li, const Mat<R, D>& u, const Mat<R, D>& ui,
                                         uint32_t creg, R* acc) {
#pragma unroll
  for (int k = 0; k < G::n; ++k) {
    if (CTRL && ((uint32_t)G::base(k) & creg) != creg) continue;
    R xr[D], xi[D], yr[D], yi[D], lr0[D], li0[D], l1r[D], l1i[D];
#pragma unroll
    for (int q = 0; q < D; ++q) {
      xr[q] = pr[G::at(k, q)];
      xi[q] = pi[G::at(k, q)];
      lr0[q] = lr[G::at(k, q)];
      li0[q] = li[G::at(k, q)];
    }
    mv<R, D>(ui, xr, xi, yr, yi);
#pragma unroll
    for (int o = 0; o < D; ++o)
#pragma unroll
      for (int c = 0; c < D; ++c) cmul_acc(acc[2 * (o * D + c)], acc[2 * (o * D + c) + 1], lr0[o], li0[o], yr[c], yi[c]);
    mvt<R, D>(u, lr0, li0, l1r, l1i);
#pragma unroll
    for (int q = 0; q < D; ++q) {
      pr[G::at(k, q)] = yr[q];
      pi[G::at(k, q)] = yi[q];
      lr[G::at(k, q)] = l1r[q];
      li[G::at(k, q)] = l1i[q];
    }
  }
}

// The backward step of a 2x2 block in SU(2) (a product of Pauli rotations
// of its target): psi <- U^{-1} psi, lambda <- U^T lambda, and instead of
// the 4 complex entries of bar = lambda_out (x) psi_in only
//   t_k = Im(lambda_in^T sigma_k psi_in) = Im Tr(bar^T U sigma_k),
// (12 FMA per pair instead of 16), which carry every angle derivative
// (d U / d theta = U X with X in su(2)).  acc = (t_x, t_y, t_z).
template <typename R, int NR, typename G, bool CTRL>
__device__ __forceinline__ void bw_su2(R* pr, R* pi, R* lr, R* li, const Mat<R, 2>& u, const Mat<R, 2>& ui,
                                       uint32_t creg, R* acc) {
#pragma unroll
  for (int k = 0; k < G::n; ++k) {
    if (CTRL && ((uint32_t)G::base(k) & creg) != creg) continue;
    R xr[2], xi[2], yr[2], yi[2], l0r[2], l0i[2], mr[2], mi[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      xr[q] = pr[G::at(k, q)];
      xi[q] = pi[G::at(k, q)];
      l0r[q] = lr[G::at(k, q)];
      l0i[q] = li[G::at(k, q)];
    }
    mv<R, 2>(ui, xr, xi, yr, yi);
    mvt<R, 2>(u, l0r, l0i, mr, mi);
    // t_x: Im(m0 y1 + m1 y0); t_y: Re(m1 y0 - m0 y1); t_z: Im(m0 y0 - m1 y1)
    acc[0] = fma(mr[0], yi[1], fma(mi[0], yr[1], fma(mr[1], yi[0], fma(mi[1], yr[0], acc[0]))));
    acc[1] = fma(mr[1], yr[0], fma(-mi[1], yi[0], fma(-mr[0], yr[1], fma(mi[0], yi[1], acc[1]))));
    acc[2] = fma(mr[0], yi[0], fma(mi[0], yr[0], fma(-mr[1], yi[1], fma(-mi[1], yr[1], acc[2]))));
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      pr[G::at(k, q)] = yr[q];
      pi[G::at(k, q)] = yi[q];
      lr[G::at(k, q)] = mr[q];
      li[G::at(k, q)] = mi[q];
    }
  }
}

// The same step in three light sweeps (fewer live registers for 4x4 blocks):
template <typename R, int NR, int D, typename G>
__device__ __forceinline__ void sweep_apply(R* re, R* im, const Mat<R, D>& m, uint32_t creg, bool transpose) {
#pragma unroll
  for (int k = 0; k < G::n; ++k) {
    if (((uint32_t)G::base(k) & creg) != creg) continue;
    R xr[D], xi[D], yr[D], yi[D];
#pragma unroll
    for (int q = 0; q < D; ++q) {
      xr[q] = re[G::at(k, q)];
      xi[q] = im[G::at(k, q)];
    }
    if (transpose)
      mvt<R, D>(m, xr, xi, yr, yi);
    else
      mv<R, D>(m, xr, xi, yr, yi);
#pragma unroll
    for (int q = 0; q < D; ++q) {
      re[G::at(k, q)] = yr[q];
      im[G::at(k, q)] = yi[q];
    }
  }
}

// 4x4 block applied in two row-halves, so only 8 matrix entries are live
// at a time (the 1024-thread backward has 128 VGPRs).  The effective matrix
// is E = M (mode 0), M^dagger (1) or M^T (2) of the pool matrix p, with the
// two target bits exchanged when `swap` (kernels use register bits A > B).
template <typename R, int NR, typename G>
__device__ __forceinline__ void sweep_apply4(R* re, R* im, cmat<R>* p, bool swap, int mode, uint32_t creg) {
  constexpr int NG = G::n;
  R tr[NG][2], ti[NG][2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    R er[2][4], ei[2][4];
#pragma unroll
    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int r = 2 * h + rr;
        const int sr = swap ? (((r & 1) << 1) | (r >> 1)) : r;
        const int sc = swap ? (((c & 1) << 1) | (c >> 1)) : c;
        const int src = mode == 0 ? sr * 4 + sc : sc * 4 + sr;
        const C2<R> v = p[src];
        er[rr][c] = v.x;
        ei[rr][c] = mode == 1 ? -v.y : v.y;
      }
#pragma unroll
    for (int k = 0; k < NG; ++k) {
      if (((uint32_t)G::base(k) & creg) != creg) continue;
      R xr[4], xi[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        xr[q] = re[G::at(k, q)];
        xi[q] = im[G::at(k, q)];
      }
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        R ar = xr[0] * er[rr][0];
        R ai = xi[0] * er[rr][0];
        ar = fma(-ei[rr][0], xi[0], ar);
        ai = fma(ei[rr][0], xr[0], ai);
#pragma unroll
        for (int c = 1; c < 4; ++c) cmul_acc(ar, ai, er[rr][c], ei[rr][c], xr[c], xi[c]);
        if (h == 0) {
          tr[k][rr] = ar;
          ti[k][rr] = ai;
        } else {
          re[G::at(k, 2 + rr)] = ar;
          im[G::at(k, 2 + rr)] = ai;
        }
      }
      if (h == 1) {
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
          re[G::at(k, rr)] = tr[k][rr];
          im[G::at(k, rr)] = ti[k][rr];
        }
      }
    }
  }
}

// acc[(o - o0) * D + c] += lambda[o] psi[c] for rows o0 .. o0 + D/2 - 1
template <typename R, int NR, int D, typename G>
__device__ __forceinline__ void sweep_grad(const R* pr, const R* pi, const R* lr, const R* li, uint32_t creg, int o0,
                                           R* acc) {
#pragma unroll
  for (int k = 0; k < G::n; ++k) {
    if (((uint32_t)G::base(k) & creg) != creg) continue;
#pragma unroll
    for (int oo = 0; oo < D / 2; ++oo) {
      const int o = o0 + oo;
#pragma unroll
      for (int c = 0; c < D; ++c)
        cmul_acc(acc[2 * (oo * D + c)], acc[2 * (oo * D + c) + 1], lr[G::at(k, o)], li[G::at(k, o)],
                 pr[G::at(k, c)], pi[G::at(k, c)]);
    }
  }
}

// In-place exchange of two registers with one v_swap_b32 per 32 bits: a
// swap inside a runtime-dispatched branch cannot be renamed away by the
// register allocator, and three moves per pair would otherwise be emitted.
__device__ __forceinline__ void vswap(float& a, float& b) { asm("v_swap_b32 %0, %1" : "+v"(a), "+v"(b)); }
__device__ __forceinline__ void vswap(double& a, double& b) {
  uint64_t x = __double_as_longlong(a), y = __double_as_longlong(b);
  uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32), yl = (uint32_t)y, yh = (uint32_t)(y >> 32);
  asm("v_swap_b32 %0, %1" : "+v"(xl), "+v"(yl));
  asm("v_swap_b32 %0, %1" : "+v"(xh), "+v"(yh));
  a = __longlong_as_double((int64_t)(((uint64_t)xh << 32) | xl));
  b = __longlong_as_double((int64_t)(((uint64_t)yh << 32) | yl));
}

template <typename R, int NR, int A>
__device__ __forceinline__ void perm_x(R* re, R* im, uint32_t creg) {
#pragma unroll
  for (int k = 0; k < NR / 2; ++k) {
    const int i = G1<NR, A>::base(k), j = i | (1 << A);
    if (((uint32_t)i & creg) != creg) continue;
    vswap(re[i], re[j]);
    vswap(im[i], im[j]);
  }
}

template <typename R, int NR, int A, int B>
__device__ __forceinline__ void perm_swap(R* re, R* im, uint32_t creg) {
#pragma unroll
  for (int k = 0; k < NR / 4; ++k) {
    const int i = G2<NR, A, B>::base(k);
    if (((uint32_t)i & creg) != creg) continue;
    const int p = i | (1 << A), q = i | (1 << B);
    vswap(re[p], re[q]);
    vswap(im[p], im[q]);
  }
}

// diagonal on one wire held by register bit A: factor d[bit]
template <typename R, int NR, int A>
__device__ __forceinline__ void diag1(R* re, R* im, R d0r, R d0i, R d1r, R d1i) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const bool hi = (i >> A) & 1;
    const R fr = hi ? d1r : d0r, fi = hi ? d1i : d0i;
    const R ar = re[i], ai = im[i];
    re[i] = fma(ar, fr, -ai * fi);
    im[i] = fma(ar, fi, ai * fr);
  }
}

template <typename R, int NR, int A>
__device__ __forceinline__ void diag1_grad(const R* pr, const R* pi, const R* lr, const R* li, R* acc) {
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const int o = (i >> A) & 1;
    cmul_acc(acc[2 * o], acc[2 * o + 1], lr[i], li[i], pr[i], pi[i]);
  }
}

// Dense 3-wire block on register bits 2..0 (wire j <-> register bit 2-j),
// evaluated through LDS so the 8x8 matrix never occupies registers.
// mode 0: M, 1: M^dagger, 2: M^T.
template <typename R, int RB>
__device__ __forceinline__ void dense3(R* re, R* im, cmat<R>* mp, int mode, const Layout& L, C2<R>* lds) {
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

// ---------------------------------------------------------------------------
// dispatch on register bit positions
// ---------------------------------------------------------------------------
#define JQ_CASE1(...)                                  \
  switch (ra) {                                        \
    case 0: { constexpr int A = 0; __VA_ARGS__; } break; \
    case 1: { constexpr int A = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int A = 2; __VA_ARGS__; } break; \
    case 3:                                            \
      if constexpr (RB > 3) { constexpr int A = RB > 3 ? 3 : 0; __VA_ARGS__; } \
      break;                                           \
    default:                                           \
      if constexpr (RB > 4) { constexpr int A = RB > 4 ? 4 : 0; __VA_ARGS__; } \
      break;                                           \
  }

#define JQ_PAIR_U(A_, B_, ...) \
  case (A_ * 8 + B_): if constexpr (A_ < RB) { constexpr int A = A_, B = B_; __VA_ARGS__; } break;

// unordered pairs, A > B
#define JQ_CASE2(...)                                                                     \
  switch (hi * 8 + lo) {                                                                  \
    JQ_PAIR_U(1, 0, __VA_ARGS__) JQ_PAIR_U(2, 0, __VA_ARGS__) JQ_PAIR_U(2, 1, __VA_ARGS__) \
    JQ_PAIR_U(3, 0, __VA_ARGS__) JQ_PAIR_U(3, 1, __VA_ARGS__) JQ_PAIR_U(3, 2, __VA_ARGS__) \
    JQ_PAIR_U(4, 0, __VA_ARGS__) JQ_PAIR_U(4, 1, __VA_ARGS__) JQ_PAIR_U(4, 2, __VA_ARGS__) \
    JQ_PAIR_U(4, 3, __VA_ARGS__)                                                          \
    default: break;                                                                       \
  }

// Diagonal index of register i: uniform part + register bits.
struct DiagLoc {
  uint32_t uni;  // contribution of thread / outside bits
  int rbit[5];   // register bit of wire j, -1 if not a register bit
  int k;
};

__device__ __forceinline__ void decode_diag_w(int k, int w0, int w1, const Layout& L, uint64_t base, DiagLoc& dl) {
  dl.k = k;
  dl.uni = 0;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    dl.rbit[j] = -1;
    if (j >= k) continue;
    const int loc = ((j < 4 ? w0 : w1) >> (8 * (j % 4))) & 0xff;
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

__device__ __forceinline__ void decode_diag(cint* g, const Layout& L, uint64_t base, DiagLoc& dl) {
  decode_diag_w(g[G_K], g[G_DW0], g[G_DW0 + 1], L, base, dl);
}

__device__ __forceinline__ uint32_t diag_index(const DiagLoc& dl, int i) {
  uint32_t x = dl.uni;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    if (j < dl.k && dl.rbit[j] >= 0) x |= (uint32_t)((i >> dl.rbit[j]) & 1) << (dl.k - 1 - j);
  return x;
}

// multiply by the diagonal d (conj when cj)
template <typename R, int RB>
__device__ __forceinline__ void apply_diag(R* re, R* im, cmat<R>* d, const DiagLoc& dl, bool cj) {
  constexpr int NR = 1 << RB;
  const R sg = cj ? R(-1) : R(1);
  if (dl.k == 1 && dl.rbit[0] >= 0) {
    const int ra = dl.rbit[0];
    const C2<R> v0 = d[0], v1 = d[1];
    JQ_CASE1(diag1<R, NR, A>(re, im, v0.x, sg * v0.y, v1.x, sg * v1.y))
    return;
  }
  bool any_reg = false;
#pragma unroll
  for (int j = 0; j < 5; ++j) any_reg = any_reg || (j < dl.k && dl.rbit[j] >= 0);
  if (!any_reg) {  // every wire on a thread or outside bit: one factor per thread
    const C2<R> v = d[dl.uni];
    const R fr = v.x, fi = sg * v.y;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const R ar = re[i], ai = im[i];
      re[i] = fma(ar, fr, -ai * fi);
      im[i] = fma(ar, fi, ai * fr);
    }
    return;
  }
  if (dl.k <= 2) {  // up to four entries, selected per register
    R dr[4], di[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const C2<R> v = d[q < (1 << dl.k) ? q : 0];
      dr[q] = v.x;
      di[q] = sg * v.y;
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
      re[i] = fma(ar, fr, -ai * fi);
      im[i] = fma(ar, fi, ai * fr);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const uint32_t x = diag_index(dl, i);
    const C2<R> v = d[x];
    const R fr = v.x, fi = sg * v.y;
    const R ar = re[i], ai = im[i];
    re[i] = fma(ar, fr, -ai * fi);
    im[i] = fma(ar, fi, ai * fr);
  }
}

template <typename R, int NR>
__device__ __forceinline__ void diag_grad(const R* pr, const R* pi, const R* lr, const R* li, const DiagLoc& dl,
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

// DPP move of a 32-bit lane value (every control used here reads a lane of
// the same row, so all lanes receive a value).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v_permlane32_swap: a' = [a.lanes 0-31, b.lanes 0-31], b' = [a.lanes 32-63, b.lanes 32-63]
__device__ __forceinline__ float swap_add32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// v_permlane16_swap: odd rows of a <-> even rows of b
__device__ __forceinline__ float swap_add16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ double swap_add32(double a, double b) {
  const uint64_t ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  return __longlong_as_double((int64_t)(((uint64_t)hi[0] << 32) | lo[0])) +
         __longlong_as_double((int64_t)(((uint64_t)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ double swap_add16(double a, double b) {
  const uint64_t ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  return __longlong_as_double((int64_t)(((uint64_t)hi[0] << 32) | lo[0])) +
         __longlong_as_double((int64_t)(((uint64_t)hi[1] << 32) | lo[1]));
}

// Transposed wave reduction of 8 per-lane values v[0..7]: halving the set
// at each exchange (lane bit 5 by permlane32 swap, bit 4 by permlane16 swap,
// bit 3 by a row_ror:8 DPP add and a select), then the sum over 8 lanes
// (quad_perm xor 1, xor 2, row_half_mirror).  Afterwards every lane L holds
// the wave total of v[L >> 3]; lanes 8k (k < nv) add it to slots[dst(k)].
// 18 VALU per 8 values instead of one full reduction per value.
template <typename R, typename Dst>
__device__ __forceinline__ void flush8(const R* v, int nv, double* slots, Dst dst) {
  R w[4], u[2];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = swap_add32(v[k], v[k + 4]);
#pragma unroll
  for (int k = 0; k < 2; ++k) u[k] = swap_add16(w[k], w[k + 2]);
  const int lane = threadIdx.x & 63;
  const R t0 = u[0] + dpp_f<0x128>(u[0]);
  const R t1 = u[1] + dpp_f<0x128>(u[1]);
  R r = (lane & 8) ? t1 : t0;
  r += dpp_f<0xB1>(r);
  r += dpp_f<0x4E>(r);
  r += dpp_f<0x141>(r);
  const int k = lane >> 3;
  if ((lane & 7) == 0 && k < nv) lds_add_lane(&slots[dst(k)], (double)r);
}

// The same for 4 values: lane bit 5 by permlane32 swaps, bit 4 by one
// permlane16 swap, then the sum over the 16 lanes of a row (row_ror:8,
// quad_perm xor 1, xor 2, row_half_mirror).  Lane 16k holds the wave total of
// v[k] and adds it to slots[dst(k)].
template <typename R, typename Dst>
__device__ __forceinline__ void flush4(const R* v, double* slots, Dst dst, int nv = 4) {
  const R w0 = swap_add32(v[0], v[2]);
  const R w1 = swap_add32(v[1], v[3]);
  R r = swap_add16(w0, w1);
  r += dpp_f<0x128>(r);
  r += dpp_f<0xB1>(r);
  r += dpp_f<0x4E>(r);
  r += dpp_f<0x141>(r);
  const int lane = threadIdx.x & 63;
  if ((lane & 15) == 0 && (lane >> 4) < nv) lds_add_lane(&slots[dst(lane >> 4)], (double)r);
}

// Reduce per-thread accumulators acc[0 .. nv) (nv <= 8) over the workgroup
// into the LDS slots base .. base + nv.
template <typename R>
__device__ __forceinline__ void flush_vals(const R* acc, int nv, double* slots, int base) {
  R v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = e < nv ? acc[e] : R(0);
  flush8<R>(v, nv, slots, [base](int k) { return base + k; });
}

struct GateCtx {
  uint32_t creg;
  bool ok;  // thread / outside control bits satisfied
};

__device__ __forceinline__ GateCtx gate_ctx(cint* g, const Layout& L, uint64_t base) {
  GateCtx c;
  c.creg = (uint32_t)g[G_CREG];
  const uint32_t cthr = (uint32_t)g[G_CTHR];
  const uint64_t cout = ((uint64_t)(uint32_t)g[G_COUT_HI] << 32) | (uint32_t)g[G_COUT_LO];
  c.ok = ((L.toff & cthr) == cthr) && ((base & cout) == cout);
  return c;
}

// Forward: apply gate record g to the register state.  FULL = false drops
// the 4x4 register blocks and the LDS dense-3 path: passes without them run
// a variant with far fewer live registers.
template <typename R, int RB, bool FULL>
__device__ __forceinline__ void fwd_gate(R* re, R* im, cint* g, cmat<R>* pool, const Layout& L,
                                         uint64_t base, C2<R>* lds) {
  constexpr int NR = 1 << RB;
  const int kind = g[G_KIND];
  cmat<R>* mp = pool + g[G_MAT];
  if constexpr (FULL) {
    if (kind == AK_DENSEK) {  // no controls; contains barriers, so all threads enter
      if constexpr (RB >= 3) dense3<R, RB>(re, im, mp, 0, L, lds);
      return;
    }
  }
  const GateCtx cx = gate_ctx(g, L, base);
  if (!cx.ok) return;
  const int ra = g[G_RA], rb = g[G_RB];
  const int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
  switch (kind) {
    case AK_MC1: {
      Mat<R, 2> m;
      load2<R>(mp, m);
      if (cx.creg == 0) {
        JQ_CASE1(fw_dense<R, NR, 2, G1<NR, A>, false>(re, im, m, 0))
      } else {
        JQ_CASE1(fw_dense<R, NR, 2, G1<NR, A>, true>(re, im, m, cx.creg))
      }
      break;
    }
    case AK_MC2: {
      if constexpr (FULL) {
        Mat<R, 4> m;
        load4<R>(mp, m, ra < rb);
        if (cx.creg == 0) {
          JQ_CASE2(fw_dense<R, NR, 4, G2<NR, A, B>, false>(re, im, m, 0))
        } else {
          JQ_CASE2(fw_dense<R, NR, 4, G2<NR, A, B>, true>(re, im, m, cx.creg))
        }
      }
      break;
    }
    case AK_MCX: {
      JQ_CASE1(perm_x<R, NR, A>(re, im, cx.creg))
      break;
    }
    case AK_MCSWAP: {
      JQ_CASE2(perm_swap<R, NR, A, B>(re, im, cx.creg))
      break;
    }
    case AK_DIAG: {
      DiagLoc dl;
      decode_diag(g, L, base, dl);
      apply_diag<R, RB>(re, im, mp, dl, false);
      break;
    }
    default:
      break;
  }
}

// Backward step of a 4x4 block in one dispatch: psi <- U^{-1} psi, the
// cotangent in two row halves (flushed per half), lambda <- U^T lambda.
// `ok` is false for threads whose thread / outside controls are not met:
// they skip the arithmetic but take part in the wave-wide flushes.
template <typename R, int NR, typename G>
__device__ __forceinline__ void bw_mc2(R* pr, R* pi, R* lr, R* li, cmat<R>* mp, cmat<R>* up, bool inv, bool sw,
                                       bool ok, uint32_t creg, double* slots, int slot) {
  if (ok) sweep_apply4<R, NR, G>(pr, pi, up, sw, inv ? 0 : 1, creg);
  if (slot >= 0) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      R acc[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = 0;
      if (ok) sweep_grad<R, NR, 4, G>(pr, pi, lr, li, creg, 2 * half, acc);
#pragma unroll
      for (int oo = 0; oo < 2; ++oo) {
        const int ko = 2 * half + oo;
        const int bo = sw ? (((ko & 1) << 1) | (ko >> 1)) : ko;
        // acc[2 (oo 4 + c) + reim] -> slot + 2 (bo 4 + s(c)) + reim (kernel row / column order)
        flush8<R>(acc + 8 * oo, 8, slots, [=](int k) {
          const int c = k >> 1;
          const int bc = sw ? (((c & 1) << 1) | (c >> 1)) : c;
          return slot + 2 * (bo * 4 + bc) + (k & 1);
        });
      }
    }
  }
  if (ok) sweep_apply4<R, NR, G>(lr, li, mp, sw, 2, creg);
}

// Backward: psi <- U^{-1} psi, bar += lambda (x) psi, lambda <- U^T lambda.
template <typename R, int RB, bool FULL>
__device__ __forceinline__ void bwd_gate(R* pr, R* pi, R* lr, R* li, cint* g, cmat<R>* pool,
                                         const Layout& L, uint64_t base, C2<R>* lds, double* slots) {
  constexpr int NR = 1 << RB;
  const int kind = g[G_KIND];
  cmat<R>* mp = pool + g[G_MAT];
  const bool inv = g[G_INV] >= 0;
  cmat<R>* ip = pool + (inv ? g[G_INV] : g[G_MAT]);
  const int slot = g[G_SLOT];
  if constexpr (FULL) {
    if (kind == AK_DENSEK) {
      if constexpr (RB >= 3) {
        dense3<R, RB>(pr, pi, inv ? ip : mp, inv ? 0 : 1, L, lds);
        dense3<R, RB>(lr, li, mp, 2, L, lds);
      }
      return;
    }
  }
  const GateCtx cx = gate_ctx(g, L, base);
  const int ra = g[G_RA], rb = g[G_RB];
  const int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
  switch (kind) {
    case AK_MC1: {
      Mat<R, 2> u, ui;
      load2<R>(mp, u);
      if (inv)
        load2<R>(ip, ui);
      else
        adjoint_of<R, 2>(u, ui);
      R acc[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) acc[e] = 0;
      if (cx.ok) {
        if (cx.creg == 0) {
          JQ_CASE1(bw_dense<R, NR, 2, G1<NR, A>, false>(pr, pi, lr, li, u, ui, 0, acc))
        } else {
          JQ_CASE1(bw_dense<R, NR, 2, G1<NR, A>, true>(pr, pi, lr, li, u, ui, cx.creg, acc))
        }
      }
      if (slot >= 0) flush_vals<R>(acc, 8, slots, slot);
      break;
    }
    case AK_MC2: {
      if constexpr (FULL) {
        if (cx.ok || slot >= 0) {
          JQ_CASE2(bw_mc2<R, NR, G2<NR, A, B>>(pr, pi, lr, li, mp, inv ? ip : mp, inv, ra < rb, cx.ok, cx.creg,
                                                slots, slot))
        }
      }
      break;
    }
    case AK_MCX: {
      if (cx.ok) {
        JQ_CASE1(perm_x<R, NR, A>(pr, pi, cx.creg); perm_x<R, NR, A>(lr, li, cx.creg))
      }
      break;
    }
    case AK_MCSWAP: {
      if (cx.ok) {
        JQ_CASE2(perm_swap<R, NR, A, B>(pr, pi, cx.creg); perm_swap<R, NR, A, B>(lr, li, cx.creg))
      }
      break;
    }
    case AK_DIAG: {
      DiagLoc dl;
      decode_diag(g, L, base, dl);
      if (inv)
        apply_diag<R, RB>(pr, pi, ip, dl, false);
      else
        apply_diag<R, RB>(pr, pi, mp, dl, true);
      if (slot >= 0) {
        R acc[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[e] = 0;
        if (dl.k == 1 && dl.rbit[0] >= 0) {
          const int ra1 = dl.rbit[0];
          const int ra = ra1;
          JQ_CASE1(diag1_grad<R, NR, A>(pr, pi, lr, li, acc))
        } else {
          diag_grad<R, NR>(pr, pi, lr, li, dl, acc);
        }
        flush_vals<R>(acc, 2 << dl.k, slots, slot);
      }
      apply_diag<R, RB>(lr, li, mp, dl, false);
      break;
    }
    default:
      break;
  }
}

// ---------------------------------------------------------------------------
// global <-> register, register <-> LDS
// ---------------------------------------------------------------------------
// Canonical layout: register bit 0 <-> tile bit 0, thread bit j <-> tile bit
// 1+j, register bits 1.. <-> the top tile bits.  The lowest tile bits are the
// lowest global bits (5 for complex64, 6 for complex128), so the lanes of a
// wave read 16-byte vectors in contiguous runs of 256 B (complex64) or 1 KiB
// (complex128) per register pair.
template <int RB, int LNT>
__device__ __forceinline__ void canonical_layout(cint* tbits, Layout& L) {
  constexpr int T = RB + LNT;
  uint64_t gthr = 0;
  uint32_t tid = threadIdx.x;
  asm volatile("" : "+v"(tid));  // see layout_lit
#pragma unroll
  for (int j = 0; j < LNT; ++j) gthr |= (uint64_t)((tid >> j) & 1u) << tbits[1 + j];
  // keep the finished offset (2 VGPRs) rather than its per-bit terms
  asm volatile("" : "+v"(gthr));
  L.toff = tid << 1;
  L.gthr = gthr;
  L.tbits = tbits;
  L.rpos[0] = 0;
#pragma unroll
  for (int b = 1; b < RB; ++b) L.rpos[b] = T - RB + b;
}

template <int RB>
__device__ __forceinline__ bool same_layout(cint* sub, const Layout& L) {
  bool same = true;
#pragma unroll
  for (int b = 0; b < RB; ++b) same = same && ((uint32_t)sub[1 + b] == L.rpos[b]);
  return same;
}

// Load/store under the canonical layout: registers (2k, 2k+1) hold adjacent
// amplitudes, moved as one 16-byte (complex64) or two 16-byte (complex128)
// vector accesses.
// |0...0>: only register 0 of the thread at global offset 0 (its register
// bits contribute nothing) holds amplitude 0.
template <typename R, int RB>
__device__ __forceinline__ void zero_state(R* re, R* im, const Layout& L, uint64_t base) {
  constexpr int NR = 1 << RB;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    re[i] = 0;
    im[i] = 0;
  }
  if ((base | L.gthr) == 0) re[0] = 1;
}

template <typename R>
__device__ __forceinline__ void ld_pair(const C2<R>* p, R& r0, R& i0, R& r1, R& i1) {
  if constexpr (sizeof(R) == 4) {
    // every tile is read once per pass: the non-temporal hint keeps the
    // streamed vectors from evicting what the pass reuses (measured: the
    // backward passes of a 24-qubit gradient 139.8 -> 136.0 us)
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    r0 = v.x;
    i0 = v.y;
    r1 = v.z;
    i1 = v.w;
  } else {
    const double2 v0 = *reinterpret_cast<const double2*>(p);
    const double2 v1 = *reinterpret_cast<const double2*>(p + 1);
    r0 = v0.x;
    i0 = v0.y;
    r1 = v1.x;
    i1 = v1.y;
  }
}

template <typename R>
__device__ __forceinline__ void st_pair(C2<R>* p, R r0, R i0, R r1, R i1) {
  if constexpr (sizeof(R) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(r0, i0, r1, i1);
  } else {
    *reinterpret_cast<double2*>(p) = make_double2(r0, i0);
    *reinterpret_cast<double2*>(p + 1) = make_double2(r1, i1);
  }
}

// Load/store under the canonical layout.  When the vector is smaller than
// a tile (phantom tile bits) the out-of-range pairs are masked; otherwise
// every access is unconditional so the loads issue back to back.
template <typename R, int RB>
__device__ __forceinline__ void load_canon(R* re, R* im, const C2<R>* src, uint64_t vec_len, const Layout& L,
                                           uint64_t base, bool full) {
  constexpr int NR = 1 << RB;
  if (full && vec_len <= (1ull << 27)) {
    // uniform pointer (scalar registers) + one 32-bit per-thread byte offset
    const uint32_t tb = (uint32_t)L.gthr * (uint32_t)sizeof(C2<R>);
#pragma unroll
    for (int i = 0; i < NR; i += 2) {
      const char* up = reinterpret_cast<const char*>(src + (base | reg_goff<RB>(L, i)));
      ld_pair<R>(reinterpret_cast<const C2<R>*>(up + tb), re[i], im[i], re[i + 1], im[i + 1]);
    }
  } else if (full) {
#pragma unroll
    for (int i = 0; i < NR; i += 2) ld_pair<R>(src + reg_gidx<RB>(L, base, i), re[i], im[i], re[i + 1], im[i + 1]);
  } else {
#pragma unroll
    for (int i = 0; i < NR; i += 2) {
      const uint64_t g = reg_gidx<RB>(L, base, i);
      const bool ok = g < vec_len;
      ld_pair<R>(src + (ok ? g : 0), re[i], im[i], re[i + 1], im[i + 1]);
      re[i] = ok ? re[i] : R(0);
      im[i] = ok ? im[i] : R(0);
      re[i + 1] = ok ? re[i + 1] : R(0);
      im[i + 1] = ok ? im[i + 1] : R(0);
    }
  }
}

// An empty asm barrier: the per-thread address base becomes opaque, so the
// compiler recomputes store addresses instead of keeping the load addresses
// of all registers live across the gate loop.
__device__ __forceinline__ uint64_t opaque(uint64_t x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <typename R, int RB>
__device__ __forceinline__ void store_canon(const R* re, const R* im, C2<R>* dst, uint64_t vec_len, const Layout& L0,
                                            uint64_t base, bool full) {
  constexpr int NR = 1 << RB;
  Layout L = L0;
  L.gthr = opaque(L0.gthr);
  if (full && vec_len <= (1ull << 27)) {
    const uint32_t tb = (uint32_t)L.gthr * (uint32_t)sizeof(C2<R>);
#pragma unroll
    for (int i = 0; i < NR; i += 2) {
      char* up = reinterpret_cast<char*>(dst + (base | reg_goff<RB>(L, i)));
      st_pair<R>(reinterpret_cast<C2<R>*>(up + tb), re[i], im[i], re[i + 1], im[i + 1]);
    }
  } else if (full) {
#pragma unroll
    for (int i = 0; i < NR; i += 2) st_pair<R>(dst + reg_gidx<RB>(L, base, i), re[i], im[i], re[i + 1], im[i + 1]);
  } else {
#pragma unroll
    for (int i = 0; i < NR; i += 2) {
      const uint64_t g = reg_gidx<RB>(L, base, i);
      if (g < vec_len) st_pair<R>(dst + g, re[i], im[i], re[i + 1], im[i + 1]);
    }
  }
}

// Register <-> LDS exchange that moves the tile from layout `from` to
// layout `to`.  LDS holds the tile as two planes (real parts, then imaginary
// parts, 2^T entries each), so re[i] and im[i] are stored from any two
// registers with one address and the DS offset field -- no register pairs
// are assembled.  The swizzle is GF(2)-linear and thread / register bits are
// disjoint, so the address of register i is
//   swz(toff) ^ XOR_{b in i} swz(1 << rpos[b])     (entries)
// whose second term is wave-uniform: it is formed in SGPRs (one s_xor per
// register) and each access costs a single v_xor.
// `region` selects a 2-plane LDS region (0: the first 2 * 2^T entries, 1:
// the next); `tail_sync` = false skips the closing barrier when the next LDS
// writes go to the other region (psi and lambda in the backward).
// The two halves of an exchange: every thread stores its registers under
// layout `from`, and (after a barrier) loads them back under layout `to`.
template <typename R, int RB, int TB = kTileBits<R>>
__device__ __forceinline__ void relayout_put(const R* re, const R* im, C2<R>* lds, const Layout& from, int region) {
  constexpr int NR = 1 << RB;
  constexpr uint32_t ES = sizeof(R);
  uint32_t wf[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) wf[b] = __builtin_amdgcn_readfirstlane(lds_swz(1u << from.rpos[b]) * ES);
  // 32-bit LDS addresses.  The tile starts the dynamic LDS block, which is
  // at LDS offset 0 (the kernels declare no static LDS), so the thread term
  // can carry the base and the register term is XORed in.
  constexpr uint32_t IM = ES << TB;  // byte distance of the imaginary plane
  const uint32_t lb = (uint32_t)(uintptr_t)(JQ_LDS char*)(char*)lds + (uint32_t)region * 2 * IM;
  const uint32_t sf = lb + lds_swz(from.toff) * ES;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < RB; ++b)
      if ((i >> b) & 1) o ^= wf[b];
    const uint32_t a = sf ^ o;
    *reinterpret_cast<JQ_LDS R*>((uintptr_t)a) = re[i];
    *reinterpret_cast<JQ_LDS R*>((uintptr_t)(a + IM)) = im[i];
  }
}

template <typename R, int RB, int TB = kTileBits<R>>
__device__ __forceinline__ void relayout_get(R* re, R* im, C2<R>* lds, const Layout& to, int region) {
  constexpr int NR = 1 << RB;
  constexpr uint32_t ES = sizeof(R);
  uint32_t wt[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) wt[b] = __builtin_amdgcn_readfirstlane(lds_swz(1u << to.rpos[b]) * ES);
  constexpr uint32_t IM = ES << TB;
  const uint32_t lb = (uint32_t)(uintptr_t)(JQ_LDS char*)(char*)lds + (uint32_t)region * 2 * IM;
  const uint32_t st = lb + lds_swz(to.toff) * ES;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < RB; ++b)
      if ((i >> b) & 1) o ^= wt[b];
    const uint32_t a = st ^ o;
    re[i] = *reinterpret_cast<const JQ_LDS R*>((uintptr_t)a);
    im[i] = *reinterpret_cast<const JQ_LDS R*>((uintptr_t)(a + IM));
  }
}

template <typename R, int RB, int TB = kTileBits<R>>
__device__ __forceinline__ void relayout(R* re, R* im, C2<R>* lds, const Layout& from, const Layout& to,
                                         int region = 0, bool tail_sync = true) {
  relayout_put<R, RB, TB>(re, im, lds, from, region);
  __syncthreads();
  relayout_get<R, RB, TB>(re, im, lds, to, region);
  if (tail_sync) __syncthreads();
}

// psi and lambda of the backward together (regions 0 and 1): both stores,
// one barrier, both loads, and (tail_sync) one barrier before the regions
// may be written again.
// With one LDS region (large tiles) psi goes through it first, then lambda.
template <typename R, int RB, int TB = kTileBits<R>>
__device__ __forceinline__ void relayout_pair(R* pr, R* pi, R* lr, R* li, C2<R>* lds, const Layout& from,
                                              const Layout& to, bool tail_sync = true) {
  if constexpr (kRegions<R, TB, RB> == 1) {
    relayout_put<R, RB, TB>(pr, pi, lds, from, 0);
    __syncthreads();
    relayout_get<R, RB, TB>(pr, pi, lds, to, 0);
    __syncthreads();
    relayout_put<R, RB, TB>(lr, li, lds, from, 0);
    __syncthreads();
    relayout_get<R, RB, TB>(lr, li, lds, to, 0);
  } else {
    relayout_put<R, RB, TB>(pr, pi, lds, from, 0);
    relayout_put<R, RB, TB>(lr, li, lds, from, 1);
    __syncthreads();
    relayout_get<R, RB, TB>(pr, pi, lds, to, 0);
    relayout_get<R, RB, TB>(lr, li, lds, to, 1);
  }
  if (tail_sync) __syncthreads();
}

__device__ __forceinline__ uint64_t tile_base(int tile, cint* tbits, int T, int n) {
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
// Observable table, staged in LDS for the per-amplitude reductions.
typedef ObsTable ObsLds;

// (need = false: only the barrier, for passes that read no observable; the
// copy's global loads would otherwise hold the tile loads behind them)
__device__ __forceinline__ void load_obs(const SvArgs& a, ObsLds* t, bool need = true) {
  const ObsTable* src = reinterpret_cast<const ObsTable*>(a.obs);
  if (need && src != nullptr) {
    for (int o = threadIdx.x; o < 64; o += blockDim.x) t->gbit[o] = src->gbit[o];
    for (int o = threadIdx.x; o < 128; o += blockDim.x) t->d[o] = src->d[o];
    for (int o = threadIdx.x; o < 64; o += blockDim.x) t->mask[o] = src->mask[o];
  }
  __syncthreads();
}

// weight of amplitude g for the lambda seed: sum_o ct_o d_o[bit]
__device__ __forceinline__ double obs_weight(uint64_t g, int n_obs, const ObsLds* t, const double* ct) {
  double w = 0;
  for (int o = 0; o < n_obs; ++o) {
    const int bit = (int)((g >> t->gbit[o]) & 1ull);
    w += ct[o] * t->d[2 * o + bit];
  }
  return w;
}

// Register-index bits whose global bit lies in mask m (wave-uniform).
template <int RB>
__device__ __forceinline__ uint32_t parity_reg_mask(const Layout& L, uint64_t m) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((m >> L.tbits[L.rpos[b]]) & 1ull) r |= 1u << b;
  return r;
}

// In-register Walsh-Hadamard transform: v[k] <- sum_i v[i] (-1)^popc(i & k).
template <typename T, int N>
__device__ __forceinline__ void wht(T* v) {
#pragma unroll
  for (int h = 1; h < N; h <<= 1)
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (!(i & h)) {
        const T x = v[i], y = v[i | h];
        v[i] = x + y;
        v[i | h] = x - y;
      }
}

// The 64-bit value v of lane l (wave-uniform l), as a wave-uniform value.
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double lane_f64(double v, int l) {
  return __longlong_as_double((long long)lane_u64((uint64_t)__double_as_longlong(v), l));
}

// Register bit holding global bit gb, or -1 (wave-uniform).
template <int RB>
__device__ __forceinline__ int reg_bit_of(const Layout& L, int gb) {
  int r = -1;
#pragma unroll
  for (int b = 0; b < RB; ++b)
    if ((int)L.tbits[L.rpos[b]] == gb) r = b;
  return r;
}

// Expectation-value partial sums of this tile (simulation.py:772-801 for
// statevectors, the diagonal trace of simulation.py:850-852 for rho).  For
// statevectors the probability mass is split once by register bit, so each
// observable costs O(1) per thread.
template <typename R, int RB>
__device__ __forceinline__ void expval_partials(const R* re, const R* im, const SvArgs& a, const ObsLds* t,
                                                const Layout& L, uint64_t base, double* ev_lds) {
  constexpr int NR = 1 << RB;
  if (a.obs_kind == JQ_OBS_SV_DIAG) {
    R ptot = 0, pb[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) pb[b] = 0;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const R p = fma(re[i], re[i], im[i] * im[i]);
      ptot += p;
#pragma unroll
      for (int b = 0; b < RB; ++b)
        if ((i >> b) & 1) pb[b] += p;
    }
    const uint64_t tg = base | L.gthr;
    for (int o = 0; o < a.n_obs; ++o) {
      const int gb = t->gbit[o];
      const R d0 = (R)t->d[2 * o], d1 = (R)t->d[2 * o + 1];
      const int rb = reg_bit_of<RB>(L, gb);
      R v;
      if (rb >= 0) {
        R p1 = 0;
#pragma unroll
        for (int b = 0; b < RB; ++b) p1 = b == rb ? pb[b] : p1;
        v = d0 * (ptot - p1) + d1 * p1;
      } else {
        v = (((tg >> gb) & 1ull) ? d1 : d0) * ptot;
      }
      lds_add(&ev_lds[o], (double)wave_sum<R>(v));
    }
    return;
  }
  if (a.obs_kind == JQ_OBS_SV_PARITY) {
    // sum_i p_i d(parity) = (d0 + d1)/2 P + (d0 - d1)/2 (-1)^pt W[mreg] with
    // W the Walsh-Hadamard transform of the thread's probabilities over the
    // register index (W[0] = P): one transform serves every Z word, whose
    // register part is then one entry (mreg is wave-uniform) and whose thread
    // and outside bits give the sign pt
    R p[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) p[i] = fma(re[i], re[i], im[i] * im[i]);
    wht<R, NR>(p);
    const uint64_t tg = base | L.gthr;  // register bit positions are zero here
    for (int o0 = 0; o0 < a.n_obs; o0 += 8) {
      const int nv = a.n_obs - o0 < 8 ? a.n_obs - o0 : 8;
      R v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        v[k] = 0;
        if (k < nv) {
          const uint64_t m = t->mask[o0 + k];
          const uint32_t mreg = __builtin_amdgcn_readfirstlane(parity_reg_mask<RB>(L, m));
          const R d0 = (R)t->d[2 * (o0 + k)], d1 = (R)t->d[2 * (o0 + k) + 1];
          const R h = (__popcll(tg & m) & 1) ? (d1 - d0) : (d0 - d1);
          v[k] = R(0.5) * fma(h, p[mreg], (d0 + d1) * p[0]);
        }
      }
      flush8<R>(v, nv, ev_lds, [o0](int k) { return o0 + k; });
    }
    return;
  }
  const int nq = a.n_wires / 2;
  const uint64_t lowmask = (1ull << nq) - 1;
  for (int o = 0; o < a.n_obs; ++o) {
    const int gb = t->gbit[o];
    const double d0 = t->d[2 * o], d1 = t->d[2 * o + 1];
    R acc = 0;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const uint64_t g = reg_gidx<RB>(L, base, i);
      const R d = ((g >> gb) & 1ull) ? (R)d1 : (R)d0;
      const R v = ((g >> nq) == (g & lowmask)) ? re[i] : R(0);
      acc = fma(v, d, acc);
    }
    lds_add(&ev_lds[o], (double)wave_sum<R>(acc));
  }
}

// ---------------------------------------------------------------------------
// building blocks of plan-specialised pass bodies
// ---------------------------------------------------------------------------
// A generated body (jit.cpp) calls these with its register bits, control
// masks, pool offsets and layouts as literals.  Gates follow one another in
// straight-line code, so a permutation of registers (CX, SWAP on register
// bits) compiles to renaming only, and no dispatch merges register states.

// Layout of a substep from literal words: RB register sources, then LNT
// thread sources (tile bits), as in the substep record.
template <int RB, int LNT, int... W>
__device__ __forceinline__ void layout_lit(cint* tbits, Layout& L) {
  constexpr int w[] = {W...};
  static_assert(sizeof...(W) == RB + LNT, "layout words");
  uint32_t toff = 0;
  uint32_t tid = threadIdx.x;
  // opaque per use: inside the persistent tile loops this keeps every
  // thread-dependent LDS address out of loop-invariant code motion, whose
  // hoisted addresses would stay live across the whole loop
  asm volatile("" : "+v"(tid));
#pragma unroll
  for (int j = 0; j < LNT; ++j) toff |= ((tid >> j) & 1u) << w[RB + j];
  L.toff = toff;
  L.gthr = 0;
  L.tbits = tbits;
#pragma unroll
  for (int b = 0; b < RB; ++b) L.rpos[b] = w[b];
}

// The same layout with its global thread offset, for the direct accesses
// below (substep layouts otherwise carry none).
template <int RB, int LNT, int... W>
__device__ __forceinline__ void layout_lit_g(cint* tbits, Layout& L) {
  constexpr int w[] = {W...};
  layout_lit<RB, LNT, W...>(tbits, L);
  uint32_t tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  uint64_t g = 0;
#pragma unroll
  for (int j = 0; j < LNT; ++j) g |= (uint64_t)((tid >> j) & 1u) << tbits[w[RB + j]];
  asm volatile("" : "+v"(g));
  L.gthr = g;
}

template <typename R>
__device__ __forceinline__ void ld_one(const C2<R>* p, R& r, R& i) {
  if constexpr (sizeof(R) == 4) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));  // read once, as ld_pair
    r = v.x;
    i = v.y;
  } else {
    const double2 v = *reinterpret_cast<const double2*>(p);
    r = v.x;
    i = v.y;
  }
}

template <typename R>
__device__ __forceinline__ void st_one(C2<R>* p, R r, R i) {
  if constexpr (sizeof(R) == 4)
    *reinterpret_cast<float2*>(p) = make_float2(r, i);
  else
    *reinterpret_cast<double2*>(p) = make_double2(r, i);
}

// Full-tile HBM access under a substep layout (with gthr) whose lanes run
// over the lowest global bits, so that every 8 lanes cover 64 contiguous
// bytes: a pass can start or end in its own substep layout instead of
// exchanging the tile through LDS into the canonical one.  PAIRS: register
// bit 0 is global bit 0, so registers (2k, 2k+1) hold adjacent amplitudes and
// move as one access.  Byte offsets are 32-bit (vectors of up to 2^27
// amplitudes); the generator selects these only for such plans.
template <typename R, int RB, bool PAIRS>
__device__ __forceinline__ void load_direct(R* re, R* im, const C2<R>* src, const Layout& L, uint64_t base) {
  constexpr int NR = 1 << RB;
  const uint32_t tb = (uint32_t)L.gthr * (uint32_t)sizeof(C2<R>);
#pragma unroll
  for (int i = 0; i < NR; i += PAIRS ? 2 : 1) {
    const char* up = reinterpret_cast<const char*>(src + (base | reg_goff<RB>(L, i))) + tb;
    if constexpr (PAIRS)
      ld_pair<R>(reinterpret_cast<const C2<R>*>(up), re[i], im[i], re[i + 1], im[i + 1]);
    else
      ld_one<R>(reinterpret_cast<const C2<R>*>(up), re[i], im[i]);
  }
}

template <typename R, int RB, bool PAIRS>
__device__ __forceinline__ void store_direct(const R* re, const R* im, C2<R>* dst, const Layout& L, uint64_t base) {
  constexpr int NR = 1 << RB;
  const uint32_t tb = (uint32_t)opaque(L.gthr) * (uint32_t)sizeof(C2<R>);
#pragma unroll
  for (int i = 0; i < NR; i += PAIRS ? 2 : 1) {
    char* up = reinterpret_cast<char*>(dst + (base | reg_goff<RB>(L, i))) + tb;
    if constexpr (PAIRS)
      st_pair<R>(reinterpret_cast<C2<R>*>(up), re[i], im[i], re[i + 1], im[i + 1]);
    else
      st_one<R>(reinterpret_cast<C2<R>*>(up), re[i], im[i]);
  }
}

template <typename R, int RB, int LNT, int... W>
__device__ __forceinline__ void jit_relayout(R* re, R* im, C2<R>* lds, Layout& L) {
  Layout nl;
  layout_lit<RB, LNT, W...>(L.tbits, nl);
  relayout<R, RB, RB + LNT>(re, im, lds, L, nl);
  L = nl;
}

template <typename R, int RB, int LNT, int... W>
__device__ __forceinline__ void jit_relayout2(R* pr, R* pi, R* lr, R* li, C2<R>* lds, Layout& L) {
  Layout nl;
  layout_lit<RB, LNT, W...>(L.tbits, nl);
  relayout_pair<R, RB, RB + LNT>(pr, pi, lr, li, lds, L, nl);
  L = nl;
}

// The last exchange of a pass body whose tile leaves for HBM directly (no
// later LDS use of the tile region before the next barrier): its closing
// barrier is dropped.
template <typename R, int RB, int LNT, int... W>
__device__ __forceinline__ void jit_relayout_nt(R* re, R* im, C2<R>* lds, Layout& L) {
  Layout nl;
  layout_lit<RB, LNT, W...>(L.tbits, nl);
  relayout<R, RB, RB + LNT>(re, im, lds, L, nl, 0, false);
  L = nl;
}

template <typename R, int RB, int LNT, int... W>
__device__ __forceinline__ void jit_relayout2_nt(R* pr, R* pi, R* lr, R* li, C2<R>* lds, Layout& L) {
  Layout nl;
  layout_lit<RB, LNT, W...>(L.tbits, nl);
  relayout_pair<R, RB, RB + LNT>(pr, pi, lr, li, lds, L, nl, false);
  L = nl;
}

// The forward's single-vector exchange with the same barrier placement
// (one tile per workgroup: the first exchange follows the tile's LDS read
// and its barrier).
template <bool LEAD, typename R, int RB, int LNT, int... W>
__device__ __forceinline__ void jit_relayout_split(R* re, R* im, C2<R>* lds, Layout& L) {
  Layout nl;
  layout_lit<RB, LNT, W...>(L.tbits, nl);
  if (LEAD) __syncthreads();
  relayout<R, RB, RB + LNT>(re, im, lds, L, nl, 0, false);
  L = nl;
}

// Exchanges of the persistent backward with the barrier that frees the LDS
// regions moved from the end of one exchange to the start of the next
// (LEAD): the waves run their gates while the slowest finishes its loads.
// The first exchange of an item needs none (the tile loop's barrier after
// the previous item's stores precedes it); writers of the regions other
// than exchanges issue their own barrier first.
template <bool LEAD, typename R, int RB, int LNT, int... W>
__device__ __forceinline__ void jit_relayout2_split(R* pr, R* pi, R* lr, R* li, C2<R>* lds, Layout& L) {
  Layout nl;
  layout_lit<RB, LNT, W...>(L.tbits, nl);
  if (LEAD) __syncthreads();
  relayout_pair<R, RB, RB + LNT>(pr, pi, lr, li, lds, L, nl, false);
  L = nl;
}

// thread / outside-tile controls of a gate
__device__ __forceinline__ bool ctl_ok(const Layout& L, uint64_t base, uint32_t cthr, uint64_t cout) {
  return ((L.toff & cthr) == cthr) && ((base & cout) == cout);
}

// X on register bit A / exchange of register bits A, B, as plain moves
template <typename R, int NR, int A, uint32_t CREG>
__device__ __forceinline__ void ren_x(R* re, R* im) {
#pragma unroll
  for (int k = 0; k < NR / 2; ++k) {
    const int i = G1<NR, A>::base(k), j = i | (1 << A);
    if (((uint32_t)i & CREG) != CREG) continue;
    const R tr = re[i], ti = im[i];
    re[i] = re[j];
    im[i] = im[j];
    re[j] = tr;
    im[j] = ti;
  }
}

template <typename R, int NR, int A, int B, uint32_t CREG>
__device__ __forceinline__ void ren_swap(R* re, R* im) {
#pragma unroll
  for (int k = 0; k < NR / 4; ++k) {
    const int i = G2<NR, A, B>::base(k);
    if (((uint32_t)i & CREG) != CREG) continue;
    const int p = i | (1 << A), q = i | (1 << B);
    const R tr = re[p], ti = im[p];
    re[p] = re[q];
    im[p] = im[q];
    re[q] = tr;
    im[q] = ti;
  }
}

// forward gates
template <typename R, int RB, int A, uint32_t CREG>
__device__ __forceinline__ void jf_mc1(R* re, R* im, cmat<R>* mp) {
  constexpr int NR = 1 << RB;
  Mat<R, 2> m;
  load2<R>(mp, m);
  fw_dense<R, NR, 2, G1<NR, A>, (CREG != 0)>(re, im, m, CREG);
}

template <typename R, int RB, int A, int B, uint32_t CREG, bool SW>
__device__ __forceinline__ void jf_mc2(R* re, R* im, cmat<R>* mp) {
  constexpr int NR = 1 << RB;
  Mat<R, 4> m;
  load4<R>(mp, m, SW);
  fw_dense<R, NR, 4, G2<NR, A, B>, (CREG != 0)>(re, im, m, CREG);
}

template <typename R, int RB, int K, int W0, int W1>
__device__ __forceinline__ void jf_diag(R* re, R* im, cmat<R>* d, const Layout& L, uint64_t base) {
  DiagLoc dl;
  decode_diag_w(K, W0, W1, L, base, dl);
  apply_diag<R, RB>(re, im, d, dl, false);
}

// backward gates: psi <- U^{-1} psi, bar += lambda (x) psi, lambda <- U^T lambda
// QUAT (blocks of Pauli rotations, in SU(2)): the cotangent is flushed as
// t_x, t_y (bar00 slot reals) and t_z (first bar10 slot real) of bw_su2;
// the gradient chain turns them back into an effective bar for U.
template <typename R, int RB, int A, uint32_t CREG, bool INV, bool QUAT = false>
__device__ __forceinline__ void jb_mc1(R* pr, R* pi, R* lr, R* li, cmat<R>* mp, cmat<R>* ip, bool ok, double* slots,
                                       int slot) {
  constexpr int NR = 1 << RB;
  Mat<R, 2> u, ui;
  load2<R>(mp, u);
  if constexpr (INV)
    load2<R>(ip, ui);
  else
    adjoint_of<R, 2>(u, ui);
  if constexpr (QUAT) {
    R acc[4] = {0, 0, 0, 0};
    if (ok) bw_su2<R, NR, G1<NR, A>, (CREG != 0)>(pr, pi, lr, li, u, ui, CREG, acc);
    if (slot < 0) return;
    flush4<R>(acc, slots, [slot](int k) { return slot + (k < 2 ? k : k + 2); }, 3);
  } else {
    R acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = 0;
    if (ok) bw_dense<R, NR, 2, G1<NR, A>, (CREG != 0)>(pr, pi, lr, li, u, ui, CREG, acc);
    if (slot < 0) return;
    flush_vals<R>(acc, 8, slots, slot);
  }
}

template <typename R, int RB, int A, int B, uint32_t CREG, bool INV, bool SW>
__device__ __forceinline__ void jb_mc2(R* pr, R* pi, R* lr, R* li, cmat<R>* mp, cmat<R>* ip, bool ok, double* slots,
                                       int slot) {
  constexpr int NR = 1 << RB;
  bw_mc2<R, NR, G2<NR, A, B>>(pr, pi, lr, li, mp, INV ? ip : mp, INV, SW, ok, CREG, slots, slot);
}

template <typename R, int RB, int K, int W0, int W1, bool INV>
__device__ __forceinline__ void jb_diag(R* pr, R* pi, R* lr, R* li, cmat<R>* mp, cmat<R>* ip, const Layout& L,
                                        uint64_t base, double* slots, int slot) {
  constexpr int NR = 1 << RB;
  DiagLoc dl;
  decode_diag_w(K, W0, W1, L, base, dl);
  if constexpr (INV)
    apply_diag<R, RB>(pr, pi, ip, dl, false);
  else
    apply_diag<R, RB>(pr, pi, mp, dl, true);
  if (slot >= 0) {
    R acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = 0;
    if (dl.k == 1 && dl.rbit[0] >= 0) {
      constexpr int A = (K == 1 && ((W0 & 0xff) >> 6) == DL_REG) ? (W0 & 63) : 0;
      diag1_grad<R, NR, A>(pr, pi, lr, li, acc);
    } else {
      diag_grad<R, NR>(pr, pi, lr, li, dl, acc);
    }
    flush_vals<R>(acc, 2 << K, slots, slot);
  }
  apply_diag<R, RB>(lr, li, mp, dl, false);
}

template <typename R, int RB, bool INV>
__device__ __forceinline__ void jb_dense3(R* pr, R* pi, R* lr, R* li, cmat<R>* mp, cmat<R>* ip, const Layout& L,
                                          C2<R>* lds) {
  dense3<R, RB>(pr, pi, INV ? ip : mp, INV ? 0 : 1, L, lds);
  dense3<R, RB>(lr, li, mp, 2, L, lds);
}

// Per-circuit arguments of grid row blockIdx.y: the launch arguments
// themselves, or (multi-plan launch) the row's group record.
template <typename R>
__device__ __forceinline__ SvArgs resolve_args(const SvArgs& a0, int& circ) {
  if (a0.groups == nullptr) {
    circ = blockIdx.y;
    return a0;
  }
  typedef const JQ_CONST SvGroup cgroup;
  cgroup* G = (cgroup*)a0.groups;
  const int row = blockIdx.y;
  int g = 0;
  for (int k = 1; k < a0.n_groups; ++k)
    if (G[k].circ_begin <= row) g = k;
  SvArgs a = a0;
  circ = row - G[g].circ_begin;
  const int P = G[g].n_passes;
  const int q = a0.dir == 0 ? a0.step : P - 1 - a0.step;
  const int64_t vec = (int64_t)a0.vec_len;
  const int64_t slice = vec * G[g].batch;  // elements of one checkpoint pass slice
  C2<R>* ck = (C2<R>*)G[g].ckpt;
  C2<R>* fin = (C2<R>*)G[g].state;
  auto pass_out = [&](int p) -> C2<R>* { return p == P - 1 ? fin : ck + slice * p; };
  a.pass = q;
  a.prog = a0.dir == 0 ? G[g].prog_f : G[g].prog_b;
  a.mats = G[g].mats;
  a.pool_stride = G[g].pool_stride;
  if (a0.dir == 0) {
    a.src = q == 0 ? G[g].psi0 : (const void*)pass_out(q - 1);
    a.src_stride = q == 0 ? G[g].psi0_stride : vec;
    a.dst = pass_out(q);
  } else {
    a.src = pass_out(q);
    a.src_stride = vec;
    a.dst = nullptr;
  }
  a.dst_stride = vec;
  a.lam = G[g].lam;
  a.lam_stride = vec;
  a.ev_part = G[g].ev_part;
  a.grad_part = G[g].grad_part;
  a.part_stride = G[g].part_stride;
  a.ct = G[g].ct;
  return a;
}

// Gate sequence of a pass.  The interpreter walks the substep and gate
// records of the program; a body generated for one plan at run time
// (jit.cpp) spells the same sequence out with literal register bits, masks
// and pool offsets, so no record is decoded and no dispatch is taken.
struct InterpBody {
  template <typename R, int RB, int NT, bool FULL>
  static __device__ __forceinline__ void fwd(cint* P, int pass_idx, R* re, R* im, Layout& L, uint64_t base,
                                             C2<R>* lds, cmat<R>* pool) {
    const int T = P[2];
    cint* pass = P + P[HDR_WORDS + pass_idx];
    const int nsub = pass[0];
    cint* tbits = pass + 4;
    cint* soff = pass + 4 + T;
    constexpr int LNT = __builtin_ctz(NT);
    for (int s = 0; s < nsub; ++s) {
      cint* sub = P + soff[s];
      const int ng = sub[0];
      if (!same_layout<RB>(sub, L)) {
        Layout nl;
        decode_layout<RB, LNT>(sub, tbits, nl);
        relayout<R, RB>(re, im, lds, L, nl);
        L = nl;
      }
      cint* g = sub + 1 + T;
      for (int q = 0; q < ng; ++q, g += GW) fwd_gate<R, RB, FULL>(re, im, g, pool, L, base, lds);
    }
  }

  template <typename R, int RB, int NT, bool FULL>
  static __device__ __forceinline__ void bwd(cint* P, int pass_idx, R* pr, R* pi, R* lr, R* li, Layout& L,
                                             uint64_t base, C2<R>* lds, cmat<R>* pool, double* slots) {
    const int T = P[2];
    cint* pass = P + P[HDR_WORDS + pass_idx];
    const int nsub = pass[0];
    cint* tbits = pass + 4;
    cint* soff = pass + 4 + T;
    constexpr int LNT = __builtin_ctz(NT);
    for (int s = nsub - 1; s >= 0; --s) {
      cint* sub = P + soff[s];
      if (!same_layout<RB>(sub, L)) {
        Layout nl;
        decode_layout<RB, LNT>(sub, tbits, nl);
        relayout_pair<R, RB>(pr, pi, lr, li, lds, L, nl);  // psi: region 0, lambda: region 1
        L = nl;
      }
      const int ng = sub[0];
      cint* g0 = sub + 1 + T;
      for (int q = ng - 1; q >= 0; --q) {
        cint* g = g0 + q * GW;
        bwd_gate<R, RB, FULL>(pr, pi, lr, li, g, pool, L, base, lds, slots);
      }
    }
  }
};

template <typename R, int RB, int NT, bool FULL, typename Body>
__device__ __forceinline__ void forward_tile(const SvArgs& a, int pass_idx, int circ, int tile, R* re, R* im,
                                             Layout& L, uint64_t& base, C2<R>* lds, bool do_load) {
  cint* P = (cint*)a.prog;
  const int T = P[2];
  cint* pass = P + P[HDR_WORDS + pass_idx];
  cint* tbits = pass + 4;
  constexpr int LNT = __builtin_ctz(NT);
  base = tile_base(tile, tbits, T, a.n_wires);
  cmat<R>* pool = (cmat<R>*)a.mats + (int64_t)circ * a.pool_stride + a.pool_off;
  canonical_layout<RB, LNT>(tbits, L);
  if (do_load) {
    if (a.src == nullptr)
      zero_state<R, RB>(re, im, L, base);
    else
      load_canon<R, RB>(re, im, reinterpret_cast<const C2<R>*>(a.src) + (int64_t)circ * a.src_stride, a.vec_len, L,
                        base, a.n_wires >= T);
  }
  Body::template fwd<R, RB, NT, FULL>(P, pass_idx, re, im, L, base, lds, pool);
}

// Bring the registers back to the canonical layout.
template <typename R, int RB, int NT>
__device__ __forceinline__ void to_canonical(R* re, R* im, cint* tbits, Layout& L, C2<R>* lds) {
  constexpr int LNT = __builtin_ctz(NT);
  Layout cl;
  canonical_layout<RB, LNT>(tbits, cl);
  bool same = true;
#pragma unroll
  for (int b = 0; b < RB; ++b) same = same && (cl.rpos[b] == L.rpos[b]);
  if (!same) relayout<R, RB, RB + LNT>(re, im, lds, L, cl);
  L = cl;  // also when no exchange was needed: substep layouts carry no gthr
}

template <typename R, int RB, int NT, bool FULL, typename Body>
__device__ __forceinline__ void sv_forward_impl(const SvArgs& a0) {
  extern __shared__ __align__(16) unsigned char smem[];
  int circ;
  const SvArgs a = resolve_args<R>(a0, circ);
  const int T = ((cint*)a.prog)[2];
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);
  ObsLds* obs = reinterpret_cast<ObsLds*>(smem + (sizeof(C2<R>) << T));
  double* ev = reinterpret_cast<double*>(obs + 1);
  const int tile = blockIdx.x;
  cint* pass = (cint*)a.prog + ((cint*)a.prog)[HDR_WORDS + a.pass];
  const bool last = (pass[1] & PF_LAST) != 0;
  const bool want_ev = last && a.obs_kind != JQ_OBS_NONE && a.ev_part != nullptr;
  if (want_ev) {
    for (int o = threadIdx.x; o < a.n_obs; o += NT) ev[o] = 0;
    load_obs(a, obs);
  }
  R re[1 << RB], im[1 << RB];
  Layout L;
  uint64_t base;
  forward_tile<R, RB, NT, FULL, Body>(a, a.pass, circ, tile, re, im, L, base, lds, true);
  to_canonical<R, RB, NT>(re, im, pass + 4, L, lds);
  store_canon<R, RB>(re, im, reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride, a.vec_len, L, base,
                     a.n_wires >= T);
  if (want_ev) {
    __syncthreads();
    expval_partials<R, RB>(re, im, a, obs, L, base, ev);
    __syncthreads();
    for (int o = threadIdx.x; o < a.n_obs; o += NT)
      a.ev_part[((int64_t)circ * a.n_obs + o) * gridDim.x + tile] = ev[o];
  }
}

template <typename R, int RB, int NT, bool FULL>
__global__ __launch_bounds__(NT, sizeof(R) == 4 ? 4 : 2) void sv_forward_kernel(SvArgs a0) {
  sv_forward_impl<R, RB, NT, FULL, InterpBody>(a0);
}

// Lambda seed: the JAX cotangent of the measurement (simulation.py:617),
// 2 conj(psi) sum_o ct_o d_o for statevectors, sum_o ct_o d_o on the
// diagonal of rho for density matrices.
template <typename R, int RB>
__device__ __forceinline__ void seed_lambda(const R* pr, const R* pi, R* lr, R* li, const SvArgs& a, const ObsLds* t,
                                            int circ, const Layout& L, uint64_t base) {
  constexpr int NR = 1 << RB;
  const double* ct = a.ct + (int64_t)circ * a.n_obs;
  // The statevector kinds stage observable o's cotangent and (d0, d1) in
  // lane o (at most 64 observables, the table's size) with one load each,
  // and read them in the loop over observables with readlane: the loop then
  // waits on no memory.
  const int lane = threadIdx.x & 63;
  const bool has = lane < a.n_obs;
  if (a.obs_kind == JQ_OBS_SV_DIAG) {
    // w(i) = c + sum_b A_b bit_b(i): observables on thread / outside bits
    // fold into c, those on register bit b into A_b.
    const uint64_t tg = base | L.gthr;
    const double ctl = has ? ct[lane] : 0.0, d0l = has ? t->d[2 * lane] : 0.0, d1l = has ? t->d[2 * lane + 1] : 0.0;
    const int gbl = has ? t->gbit[lane] : 0;
    double c = 0, A[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) A[b] = 0;
    for (int o = 0; o < a.n_obs; ++o) {
      const int gb = __builtin_amdgcn_readlane(gbl, o);
      const double d0 = lane_f64(d0l, o), d1 = lane_f64(d1l, o), w = lane_f64(ctl, o);
      const int rb = reg_bit_of<RB>(L, gb);
      if (rb >= 0) {
        c += w * d0;
#pragma unroll
        for (int b = 0; b < RB; ++b)
          if (b == rb) A[b] += w * (d1 - d0);
      } else {
        c += w * (((tg >> gb) & 1ull) ? d1 : d0);
      }
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      double w = c;
#pragma unroll
      for (int b = 0; b < RB; ++b)
        if ((i >> b) & 1) w += A[b];
      const R w2 = (R)(2.0 * w);
      lr[i] = w2 * pr[i];
      li[i] = -w2 * pi[i];
    }
    return;
  }
  if (a.obs_kind == JQ_OBS_SV_PARITY) {
    // w_i = sum_o ct_o d_o[parity] = C + sum_o h_o (-1)^(pt_o + popc(i & mreg_o)):
    // each Z word adds its half-sum to C and its half-difference h_o, signed
    // by its thread / outside bits (pt_o), to every register's weight with
    // the wave-uniform sign of its register mask: one FMA per register.  A
    // word with no register bit (mreg == 0, a uniform branch) adds h to a
    // per-thread constant instead.  The sign bits (-1)^popc(i & mreg) are
    // built by doubling over i, one scalar XOR per register.
    const uint64_t tg = base | L.gthr;
    const double ctl = has ? ct[lane] : 0.0;
    const double c0l = has ? ctl * t->d[2 * lane] : 0.0, c1l = has ? ctl * t->d[2 * lane + 1] : 0.0;
    const double hsl = 0.5 * (c0l + c1l), hdl = 0.5 * (c0l - c1l);
    const uint64_t ml = has ? t->mask[lane] : 0;
    const uint32_t mregl = parity_reg_mask<RB>(L, ml);
    double w[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) w[i] = 0;
    double C = 0, Ct = 0;
    for (int o = 0; o < a.n_obs; ++o) {
      const uint64_t m = lane_u64(ml, o);
      const uint32_t mreg = (uint32_t)__builtin_amdgcn_readlane((int)mregl, o);
      C += lane_f64(hsl, o);
      const double hd = lane_f64(hdl, o);
      const double h = (__popcll(tg & m) & 1) ? -hd : hd;
      if (mreg == 0) {
        Ct += h;
        continue;
      }
      uint32_t sg[NR];
      sg[0] = 0;
#pragma unroll
      for (int i = 1; i < NR; ++i) sg[i] = sg[i & (i - 1)] ^ (((mreg >> __builtin_ctz(i)) & 1u) << 31);
#pragma unroll
      for (int i = 0; i < NR; ++i)
        w[i] = fma(__longlong_as_double((long long)(0x3FF0000000000000ull | ((uint64_t)sg[i] << 32))), h, w[i]);
    }
    C += Ct;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const R w2 = (R)(2.0 * (C + w[i]));
      lr[i] = w2 * pr[i];
      li[i] = -w2 * pi[i];
    }
    return;
  }
  const int nq = a.n_wires / 2;
  const uint64_t lowmask = (1ull << nq) - 1;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const uint64_t g = reg_gidx<RB>(L, base, i);
    const double w = obs_weight(g, a.n_obs, t, ct);
    lr[i] = ((g >> nq) == (g & lowmask) && g < a.vec_len) ? (R)w : R(0);
    li[i] = 0;
  }
}

template <typename R, int RB, int NT>
__device__ __forceinline__ void to_canonical2(R* pr, R* pi, R* lr, R* li, cint* tbits, Layout& L, C2<R>* lds) {
  constexpr int LNT = __builtin_ctz(NT);
  Layout cl;
  canonical_layout<RB, LNT>(tbits, cl);
  bool same = true;
#pragma unroll
  for (int b = 0; b < RB; ++b) same = same && (cl.rpos[b] == L.rpos[b]);
  if (!same) {
    relayout_pair<R, RB, RB + LNT>(pr, pi, lr, li, lds, L, cl);
  }
  L = cl;  // also when no exchange was needed: substep layouts carry no gthr
}

// Backward sweep over one pass (reverse substeps, reverse gates).  When
// `loaded` is false, psi (the pass output) and lambda are read under the
// canonical layout; otherwise they arrive in registers under layout L.
template <typename R, int RB, int NT, bool FULL, typename Body>
__device__ __forceinline__ void backward_tile(const SvArgs& a, int pass_idx, int circ, int tile, R* pr, R* pi, R* lr,
                                              R* li, Layout& L, uint64_t& base, C2<R>* lds, double* slots,
                                              const ObsLds* obs, bool loaded) {
  cint* P = (cint*)a.prog;
  const int T = P[2];
  cint* pass = P + P[HDR_WORDS + pass_idx];
  const int flags = pass[1];
  cint* tbits = pass + 4;
  constexpr int LNT = __builtin_ctz(NT);
  base = tile_base(tile, tbits, T, a.n_wires);
  cmat<R>* pool = (cmat<R>*)a.mats + (int64_t)circ * a.pool_stride + a.pool_off;
  if (!loaded) {
    canonical_layout<RB, LNT>(tbits, L);
    const C2<R>* ps = reinterpret_cast<const C2<R>*>(a.src) + (int64_t)circ * a.src_stride;
    load_canon<R, RB>(pr, pi, ps, a.vec_len, L, base, a.n_wires >= T);
    if ((flags & PF_LAST) && !a.lam_given) {
      seed_lambda<R, RB>(pr, pi, lr, li, a, obs, circ, L, base);
    } else {
      const C2<R>* ls = reinterpret_cast<const C2<R>*>(a.lam) + (int64_t)circ * a.lam_stride;
      load_canon<R, RB>(lr, li, ls, a.vec_len, L, base, a.n_wires >= T);
    }
  }
  Body::template bwd<R, RB, NT, FULL>(P, pass_idx, pr, pi, lr, li, L, base, lds, pool, slots);
  to_canonical2<R, RB, NT>(pr, pi, lr, li, tbits, L, lds);
}

template <typename R, int RB, int NT, bool FULL, typename Body>
__device__ __forceinline__ void sv_backward_impl(const SvArgs& a0) {
  extern __shared__ __align__(16) unsigned char smem[];
  int circ;
  const SvArgs a = resolve_args<R>(a0, circ);
  const int T = ((cint*)a.prog)[2];
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);  // psi and lambda tiles (relayout2)
  ObsLds* obs = reinterpret_cast<ObsLds*>(smem + ((size_t)lds_regions<R, RB>(T) * sizeof(C2<R>) << T));
  double* slots = reinterpret_cast<double*>(obs + 1);
  const int tile = blockIdx.x;
  cint* pass = (cint*)a.prog + ((cint*)a.prog)[HDR_WORDS + a.pass];
  const int nslots = pass[2];
  const int slot_base = pass[3];
  for (int e = threadIdx.x; e < nslots; e += NT) slots[e] = 0;
  load_obs(a, obs);
  R pr[1 << RB], pi[1 << RB], lr[1 << RB], li[1 << RB];
  Layout L;
  uint64_t base;
  backward_tile<R, RB, NT, FULL, Body>(a, a.pass, circ, tile, pr, pi, lr, li, L, base, lds, slots, obs, false);
  C2<R>* lam = reinterpret_cast<C2<R>*>(a.lam) + (int64_t)circ * a.lam_stride;
  store_canon<R, RB>(lr, li, lam, a.vec_len, L, base, a.n_wires >= T);
  if (a.dst != nullptr)
    store_canon<R, RB>(pr, pi, reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride, a.vec_len, L, base, a.n_wires >= T);
  __syncthreads();
  // partials are real-major ([real][tile]) so the reduction reads them contiguously
  double* gp = a.grad_part + (int64_t)circ * a.part_stride + slot_base + tile;
  for (int e = threadIdx.x; e < nslots; e += NT) gp[(int64_t)e * gridDim.x] = slots[e];
}

template <typename R, int RB, int NT, bool FULL>
__global__ __launch_bounds__(NT) void sv_backward_kernel(SvArgs a0) {
  sv_backward_impl<R, RB, NT, FULL, InterpBody>(a0);
}

// ---------------------------------------------------------------------------
// persistent pass kernels
// ---------------------------------------------------------------------------
// A workgroup walks the work items blockIdx.x, blockIdx.x + gridDim.x, ...
// (single-plan launches: (circuit, tile) pairs of all rows).  The HBM loads of the next tile are
// issued into a second register set before the current tile's gates run;
// barriers inside a pass wait on LDS only, so the loads stay in flight
// behind the gate arithmetic and the relayouts.  Results and partial sums
// are those of the one-tile-per-workgroup kernels.
// Forward of one tile per workgroup whose tile streams into LDS
// (global_load_lds: HBM -> LDS without registers).  Each lane of the
// canonical layout moves its 16-byte amplitude pair to LDS byte
// 16 * (pair index * NT + thread), which is byte 8 * x of the amplitude at
// tile offset x: the tile lies in natural order, so a thread reads its
// amplitudes under ANY layout with one 8-byte LDS read each -- the pass
// starts directly in its first substep's layout (Body::first_layout,
// Body::fwd0) instead of loading canonically and re-laying out.  The tile
// lands in the re-layout region itself (read out before the first
// re-layout), so a workgroup needs one tile of LDS and two fit per CU.
// Complex64 with whole tiles only (vector >= tile).
template <typename R, int RB, int NT, typename Body>
__device__ __forceinline__ void sv_forward_dma(const SvArgs& a) {
  static_assert(sizeof(R) == 4, "the natural-order prefetch tile is laid out for complex64");
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NR = 1 << RB;
  constexpr int LNT = __builtin_ctz(NT);
  const int circ = blockIdx.y;
  cint* P = (cint*)a.prog;
  const int T = P[2];
  const int ntiles = P[8];
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);
  unsigned char* pre = smem;
  ObsLds* obs = reinterpret_cast<ObsLds*>(pre + (sizeof(C2<R>) << T));
  double* ev = reinterpret_cast<double*>(obs + 1);
  cint* pass = P + P[HDR_WORDS + a.pass];
  cint* tbits = pass + 4;
  const bool last = (pass[1] & PF_LAST) != 0;
  const bool want_ev = last && a.obs_kind != JQ_OBS_NONE && a.ev_part != nullptr;
  const int tile = blockIdx.x;
  if (tile >= ntiles) return;
  if (want_ev) {
    for (int o = threadIdx.x; o < a.n_obs; o += NT) ev[o] = 0;
    load_obs(a, obs);
  }
  cmat<R>* pool = (cmat<R>*)a.mats + (int64_t)circ * a.pool_stride + a.pool_off;
  const bool full = Body::kFull;  // n_wires >= T, fixed when the kernel is generated
  const bool from_zero = a.src == nullptr;
  const bool dma = full && !from_zero && a.vec_len <= (1ull << 27);  // 32-bit per-thread byte offsets
  const C2<R>* src = from_zero ? nullptr : reinterpret_cast<const C2<R>*>(a.src) + (int64_t)circ * a.src_stride;
  C2<R>* dst = reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride;
  // issue the LDS loads of the tile (canonical layout: pair r of thread t)
  auto prefetch = [&](uint64_t b) {
    Layout Lc;
    canonical_layout<RB, LNT>(tbits, Lc);
    const uint32_t tb = (uint32_t)Lc.gthr * (uint32_t)sizeof(C2<R>);
    const uint32_t wave_off = (threadIdx.x & ~63u) * 16u;
#pragma unroll
    for (int r = 0; r < NR / 2; ++r) {
      const char* up = reinterpret_cast<const char*>(src + (b | reg_goff<RB>(Lc, 2 * r))) + tb;
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)up,
                                       (JQ_LDS void*)(JQ_LDS char*)(char*)(pre + (uint32_t)r * NT * 16u + wave_off),
                                       16, 0, 0);
    }
  };
  const uint64_t base = tile_base(tile, tbits, T, a.n_wires);
  // Body::kLoad: the first substep's lanes run over the lowest global bits,
  // so the tile is read straight into registers under that layout (no LDS
  // round trip)
  const bool direct = Body::kLoad != 0 && dma;
  if (dma && !direct) prefetch(base);
  {
    R re[NR], im[NR];
    Layout L;
    bool started = false;  // registers already hold the first substep's layout
    if (direct) {
      if constexpr (Body::kLoad != 0) {
        Body::template first_layout_g<RB, LNT>(tbits, L);
        load_direct<R, RB, Body::kLoad == 1>(re, im, src, L, base);
      }
      started = true;
    } else if (dma) {
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's LDS loads have landed
      __syncthreads();                     // ... and every other wave's
      Body::template first_layout<RB, LNT>(tbits, L);
      const uint32_t lb = (uint32_t)(uintptr_t)(JQ_LDS char*)(char*)pre + L.toff * 8u;
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        const float2 v = *reinterpret_cast<const JQ_LDS float2*>((uintptr_t)(lb + reg_roff<RB>(L, i) * 8u));
        re[i] = v.x;
        im[i] = v.y;
      }
      __syncthreads();  // the prefetch tile is free again
      started = true;
    } else {
      canonical_layout<RB, LNT>(tbits, L);
      if (from_zero)
        zero_state<R, RB>(re, im, L, base);
      else
        load_canon<R, RB>(re, im, src, a.vec_len, L, base, full);
    }
    if (started)
      Body::template fwd0<R, RB, NT, false>(P, a.pass, re, im, L, base, lds, pool);
    else
      Body::template fwd<R, RB, NT, false>(P, a.pass, re, im, L, base, lds, pool);
    if constexpr (Body::kStore != 0) {
      Body::template last_layout<RB, LNT>(tbits, L);
      store_direct<R, RB, Body::kStore == 1>(re, im, dst, L, base);
    } else {
      if constexpr (Body::kSplit) __syncthreads();  // the body's last exchange has no closing barrier
      to_canonical<R, RB, NT>(re, im, tbits, L, lds);
      store_canon<R, RB>(re, im, dst, a.vec_len, L, base, full);
    }
    if (want_ev) {
      __syncthreads();
      expval_partials<R, RB>(re, im, a, obs, L, base, ev);
      __syncthreads();
      for (int o = threadIdx.x; o < a.n_obs; o += NT) {
        a.ev_part[((int64_t)circ * a.n_obs + o) * ntiles + tile] = ev[o];
        ev[o] = 0;
      }
    }
  }
}

template <typename R, int RB, int NT, typename Body>
__device__ __forceinline__ void sv_backward_loop(const SvArgs& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NR = 1 << RB;
  constexpr int LNT = __builtin_ctz(NT);
  cint* P = (cint*)a.prog;
  const int T = P[2];
  const int ntiles = P[8];
  const int tlog = __builtin_ctz((unsigned)ntiles);
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);  // psi and lambda tiles (relayout2)
  ObsLds* obs = reinterpret_cast<ObsLds*>(smem + ((size_t)lds_regions<R, RB>(T) * sizeof(C2<R>) << T));
  // two cotangent slot buffers, alternating from item to item: a buffer is
  // read (and cleared) right after its item's flushes and next written two
  // items later, so no barrier is needed to clear it
  double* slot_buf = reinterpret_cast<double*>(obs + 1);
  cint* pass = P + P[HDR_WORDS + a.pass];
  cint* tbits = pass + 4;
  const bool seed = (pass[1] & PF_LAST) != 0 && !a.lam_given;
  const int nslots = pass[2];
  const int slot_base = pass[3];
  // work items w = circuit * ntiles + tile over all rows of the launch, so
  // the resident workgroups balance across circuits when a row has few tiles
  const int64_t total = (int64_t)a.rows << tlog;
  int64_t w = blockIdx.x;
  if (w >= total) return;
  for (int e = threadIdx.x; e < 2 * nslots; e += NT) slot_buf[e] = 0;
  load_obs(a, obs, seed);  // (ends with a barrier)
  const bool full = Body::kFull;  // n_wires >= T, fixed when the kernel is generated
  const C2<R>* src = reinterpret_cast<const C2<R>*>(a.src);
  C2<R>* lam = reinterpret_cast<C2<R>*>(a.lam);
  // Tiles are read under the canonical layout, or (Body::kLoad) directly
  // under the layout of the first substep the sweep visits.
  Layout Lc;
  canonical_layout<RB, LNT>(tbits, Lc);
  // (the first layout is rebuilt per use from the opaque thread index, so
  // its offsets are not kept live across the tile loop)
  auto fetch = [&](R* xr, R* xi, const C2<R>* p, uint64_t b) {
    if constexpr (Body::kLoad != 0) {
      Layout Lf;
      Body::template first_layout<RB, LNT>(tbits, Lf);
      load_direct<R, RB, Body::kLoad == 1>(xr, xi, p, Lf, b);
    } else {
      load_canon<R, RB>(xr, xi, p, a.vec_len, Lc, b, full);
    }
  };
  R npr[NR], npi[NR], nlr[NR], nli[NR];
  int ncirc = (int)(w >> tlog), ntile = (int)(w & (ntiles - 1));
  uint64_t nbase = tile_base(ntile, tbits, T, a.n_wires);
  fetch(npr, npi, src + (int64_t)ncirc * a.src_stride, nbase);
  if (!seed) fetch(nlr, nli, lam + (int64_t)ncirc * a.lam_stride, nbase);
  int par = 0;
  for (;;) {
    const int circ = ncirc, tile = ntile;
    double* slots = slot_buf + par * nslots;
    const uint64_t base = nbase;
    R pr[NR], pi[NR], lr[NR], li[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      pr[i] = npr[i];
      pi[i] = npi[i];
    }
    if (!seed) {
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        lr[i] = nlr[i];
        li[i] = nli[i];
      }
    }
    const int64_t next = w + gridDim.x;
    if (next < total) {
      ncirc = (int)(next >> tlog);
      ntile = (int)(next & (ntiles - 1));
      nbase = tile_base(ntile, tbits, T, a.n_wires);
      fetch(npr, npi, src + (int64_t)ncirc * a.src_stride, nbase);
      if (!seed) fetch(nlr, nli, lam + (int64_t)ncirc * a.lam_stride, nbase);
    }
    cmat<R>* pool = (cmat<R>*)a.mats + (int64_t)circ * a.pool_stride + a.pool_off;
    Layout L;
    if constexpr (Body::kLoad != 0)
      Body::template first_layout<RB, LNT>(tbits, L);
    else
      canonical_layout<RB, LNT>(tbits, L);
    if (seed) seed_lambda<R, RB>(pr, pi, lr, li, a, obs, circ, L, base);
    if constexpr (Body::kLoad != 0)
      Body::template bwd0<R, RB, NT, false>(P, a.pass, pr, pi, lr, li, L, base, lds, pool, slots);
    else
      Body::template bwd<R, RB, NT, false>(P, a.pass, pr, pi, lr, li, L, base, lds, pool, slots);
    C2<R>* pout = a.dst != nullptr ? reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride : nullptr;
    if constexpr (Body::kStore != 0) {
      Body::template last_layout<RB, LNT>(tbits, L);
      store_direct<R, RB, Body::kStore == 1>(lr, li, lam + (int64_t)circ * a.lam_stride, L, base);
      if (pout != nullptr) store_direct<R, RB, Body::kStore == 1>(pr, pi, pout, L, base);
    } else if (pout != nullptr) {
      if constexpr (Body::kSplit) __syncthreads();  // the body's last exchange has no closing barrier
      to_canonical2<R, RB, NT>(pr, pi, lr, li, tbits, L, lds);
      store_canon<R, RB>(lr, li, lam + (int64_t)circ * a.lam_stride, a.vec_len, L, base, full);
      store_canon<R, RB>(pr, pi, pout, a.vec_len, L, base, full);
    } else {
      // psi is not kept: only lambda goes back to the canonical layout
      if constexpr (Body::kSplit) __syncthreads();  // the body's last exchange has no closing barrier
      to_canonical<R, RB, NT>(lr, li, tbits, L, lds);
      store_canon<R, RB>(lr, li, lam + (int64_t)circ * a.lam_stride, a.vec_len, L, base, full);
    }
    __syncthreads();  // every flush of this item has landed
    // partials are real-major ([real][tile]) so the reduction reads them contiguously
    double* gp = a.grad_part + (int64_t)circ * a.part_stride + slot_base + tile;
    for (int e = threadIdx.x; e < nslots; e += NT) {
      gp[(int64_t)e * ntiles] = slots[e];
      slots[e] = 0;
    }
    if (next >= total) break;
    w = next;
    par ^= 1;
  }
}

// Backward of one tile per workgroup for plan-specialised bodies (workgroups
// of fewer than 512 threads, or 14-bit tiles): no next-tile prefetch, so the
// state registers fit the budget of two workgroups per CU; psi and lambda
// are read under the first substep's layout and lambda (and psi) written
// under the last one when the body allows it (Body::kLoad / kStore), as the
// persistent loop does.
template <typename R, int RB, int NT, typename Body>
__device__ __forceinline__ void sv_backward_direct(const SvArgs& a0) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NR = 1 << RB;
  constexpr int LNT = __builtin_ctz(NT);
  int circ;
  const SvArgs a = resolve_args<R>(a0, circ);
  cint* P = (cint*)a.prog;
  const int T = P[2];
  C2<R>* lds = reinterpret_cast<C2<R>*>(smem);
  ObsLds* obs = reinterpret_cast<ObsLds*>(smem + ((size_t)lds_regions<R, RB>(T) * sizeof(C2<R>) << T));
  double* slots = reinterpret_cast<double*>(obs + 1);
  cint* pass = P + P[HDR_WORDS + a.pass];
  cint* tbits = pass + 4;
  const bool seed = (pass[1] & PF_LAST) != 0 && !a.lam_given;
  const int nslots = pass[2];
  const int slot_base = pass[3];
  const int tile = blockIdx.x;
  for (int e = threadIdx.x; e < nslots; e += NT) slots[e] = 0;
  load_obs(a, obs, seed);  // (ends with a barrier)
  const bool full = Body::kFull;
  const uint64_t base = tile_base(tile, tbits, T, a.n_wires);
  const C2<R>* src = reinterpret_cast<const C2<R>*>(a.src) + (int64_t)circ * a.src_stride;
  C2<R>* lam = reinterpret_cast<C2<R>*>(a.lam) + (int64_t)circ * a.lam_stride;
  R pr[NR], pi[NR], lr[NR], li[NR];
  Layout L;
  if constexpr (Body::kLoad != 0) {
    Body::template first_layout<RB, LNT>(tbits, L);
    load_direct<R, RB, Body::kLoad == 1>(pr, pi, src, L, base);
    if (!seed) load_direct<R, RB, Body::kLoad == 1>(lr, li, lam, L, base);
  } else {
    canonical_layout<RB, LNT>(tbits, L);
    load_canon<R, RB>(pr, pi, src, a.vec_len, L, base, full);
    if (!seed) load_canon<R, RB>(lr, li, lam, a.vec_len, L, base, full);
  }
  cmat<R>* pool = (cmat<R>*)a.mats + (int64_t)circ * a.pool_stride + a.pool_off;
  if (seed) seed_lambda<R, RB>(pr, pi, lr, li, a, obs, circ, L, base);
  if constexpr (Body::kLoad != 0)
    Body::template bwd0<R, RB, NT, false>(P, a.pass, pr, pi, lr, li, L, base, lds, pool, slots);
  else
    Body::template bwd<R, RB, NT, false>(P, a.pass, pr, pi, lr, li, L, base, lds, pool, slots);
  C2<R>* pout = a.dst != nullptr ? reinterpret_cast<C2<R>*>(a.dst) + (int64_t)circ * a.dst_stride : nullptr;
  if constexpr (Body::kStore != 0) {
    Body::template last_layout<RB, LNT>(tbits, L);
    store_direct<R, RB, Body::kStore == 1>(lr, li, lam, L, base);
    if (pout != nullptr) store_direct<R, RB, Body::kStore == 1>(pr, pi, pout, L, base);
  } else if (pout != nullptr) {
    to_canonical2<R, RB, NT>(pr, pi, lr, li, tbits, L, lds);
    store_canon<R, RB>(lr, li, lam, a.vec_len, L, base, full);
    store_canon<R, RB>(pr, pi, pout, a.vec_len, L, base, full);
  } else {
    to_canonical<R, RB, NT>(lr, li, tbits, L, lds);
    store_canon<R, RB>(lr, li, lam, a.vec_len, L, base, full);
  }
  __syncthreads();  // every flush has landed
  double* gp = a.grad_part + (int64_t)circ * a.part_stride + slot_base + tile;
  for (int e = threadIdx.x; e < nslots; e += NT) gp[(int64_t)e * gridDim.x] = slots[e];
}

#ifndef __HIPCC_RTC__
// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename R, int RB, int NT, bool FULL>
hipError_t launch_forward(const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  (void)hipFuncSetAttribute((const void*)sv_forward_kernel<R, RB, NT, FULL>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL((sv_forward_kernel<R, RB, NT, FULL>), dim3(ntiles, batch), dim3(NT), smem, st, a);
  return hipGetLastError();
}

template <typename R, int RB, int NT, bool FULL>
hipError_t launch_backward(const SvArgs& a, int ntiles, int batch, size_t smem, hipStream_t st) {
  (void)hipFuncSetAttribute((const void*)sv_backward_kernel<R, RB, NT, FULL>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL((sv_backward_kernel<R, RB, NT, FULL>), dim3(ntiles, batch), dim3(NT), smem, st, a);
  return hipGetLastError();
}

#endif

}  // namespace jq

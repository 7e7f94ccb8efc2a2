// Batched propagators of time-dependent Hamiltonians H(t) = sum_k f_k(t) H_k
// and their vector-Jacobian product, in FP64.
//
// The reference solves dU/dt = -i H(t) U per pulse gate with a fixed-step
// commutator-free Magnus scheme (jaqsi/evolution.py:194-263: Magnus-2 one
// exponential per step, CFM4 two), adaptive Dormand-Prince (265-351) or, for
// a single term, the closed form exp(-i F H) with F the pulse area
// (353-404), and differentiates through it with autodiff.  Every one of
// those schemes is an ordered product of exponentials
//     U = exp(-i A_{M-1}) ... exp(-i A_1) exp(-i A_0),   A_m = sum_k a[m][k] H_k
// with real coefficients a (the time step times the coefficient samples,
// combined per scheme on the host), so one kernel serves them all:
//  - one wavefront per gate; lane l owns the contiguous factors
//    [l M / 64, (l + 1) M / 64) and multiplies them in registers;
//  - the 64 lane products are combined in LDS (each lane forms its own
//    exclusive prefix and suffix by ordered products over the others);
//  - exp(X), X = -i A, is a degree-16 Taylor polynomial in Horner form after
//    scaling by 2^-s (||X||_1 / 2^s <= 1/4, truncation < 1e-22) and s
//    squarings;
//  - the VJP of a real loss L with gradient G = dL/dRe U + i dL/dIm U:
//    dL/da[m][k] = Re Tr(M_m^dag dE_m[-i H_k]) = Im Tr(H_k Z_m), with
//    M_m = L_m^dag G R_m^dag (L_m, R_m the products after / before factor m)
//    and Z_m = dE_m[M_m^dag] the Frechet derivative of exp at X_m, evaluated
//    by differentiating the same Horner recursion and squarings.  Factors
//    are unitary, so the sweep recovers R_m = E_m^dag R_{m+1} and
//    L_{m-1}^dag G = E_m^dag L_m^dag G without storing intermediates.
#include <hip/hip_runtime.h>

#include "evolve.h"
#include "jaqsi_amd.h"

namespace jq {

namespace {

struct cx {
  double x, y;
};
__device__ __forceinline__ cx cmulx(cx a, cx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cx caddx(cx a, cx b) { return {a.x + b.x, a.y + b.y}; }

template <int D>
struct Mx {
  cx e[D * D];
};

template <int D>
__device__ __forceinline__ void mx_id(Mx<D>& m) {
#pragma unroll
  for (int i = 0; i < D * D; ++i) m.e[i] = {(i / D) == (i % D) ? 1.0 : 0.0, 0.0};
}

template <int D>
__device__ __forceinline__ void mx_mul(const Mx<D>& a, const Mx<D>& b, Mx<D>& c) {  // c = a b
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int q = 0; q < D; ++q) {
      cx s = {0.0, 0.0};
#pragma unroll
      for (int k = 0; k < D; ++k) s = caddx(s, cmulx(a.e[r * D + k], b.e[k * D + q]));
      c.e[r * D + q] = s;
    }
}

template <int D>
__device__ __forceinline__ void mx_adj(const Mx<D>& a, Mx<D>& o) {  // o = a^dagger
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int q = 0; q < D; ++q) o.e[r * D + q] = {a.e[q * D + r].x, -a.e[q * D + r].y};
}

constexpr int TAYLOR = 16;

// X = -i sum_k a_k H_k
template <int D>
__device__ __forceinline__ void make_x(const double* a, const cx* H, int K, Mx<D>& X) {
#pragma unroll
  for (int i = 0; i < D * D; ++i) X.e[i] = {0.0, 0.0};
  for (int k = 0; k < K; ++k) {
    const double c = a[k];
#pragma unroll
    for (int i = 0; i < D * D; ++i) {
      const cx h = H[k * D * D + i];
      X.e[i].x += c * h.y;  // -i c h
      X.e[i].y -= c * h.x;
    }
  }
}

template <int D>
__device__ __forceinline__ int scaling(const Mx<D>& X) {
  double n1 = 0;  // max column sum
#pragma unroll
  for (int q = 0; q < D; ++q) {
    double s = 0;
#pragma unroll
    for (int r = 0; r < D; ++r) s += fabs(X.e[r * D + q].x) + fabs(X.e[r * D + q].y);
    n1 = fmax(n1, s);
  }
  int s = 0;
  while (n1 > 0.25 && s < 60) {
    n1 *= 0.5;
    ++s;
  }
  return s;
}

// E = exp(X)
template <int D>
__device__ __forceinline__ void expm(const Mx<D>& X0, Mx<D>& E) {
  const int s = scaling<D>(X0);
  const double sc = ldexp(1.0, -s);
  Mx<D> X, T;
#pragma unroll
  for (int i = 0; i < D * D; ++i) X.e[i] = {X0.e[i].x * sc, X0.e[i].y * sc};
  mx_id<D>(E);  // Horner: P_N = I, P_m = I + X P_{m+1} / (m + 1)
  for (int m = TAYLOR - 1; m >= 0; --m) {
    mx_mul<D>(X, E, T);
    const double inv = 1.0 / (m + 1);
#pragma unroll
    for (int i = 0; i < D * D; ++i) E.e[i] = {((i / D) == (i % D) ? 1.0 : 0.0) + T.e[i].x * inv, T.e[i].y * inv};
  }
  for (int k = 0; k < s; ++k) {
    mx_mul<D>(E, E, T);
    E = T;
  }
}

// E = exp(X) and F = dE[Y] (Frechet derivative in direction Y)
template <int D>
__device__ __forceinline__ void expm_frechet(const Mx<D>& X0, const Mx<D>& Y0, Mx<D>& E, Mx<D>& F) {
  const int s = scaling<D>(X0);
  const double sc = ldexp(1.0, -s);
  Mx<D> X, Y, T, U;
#pragma unroll
  for (int i = 0; i < D * D; ++i) {
    X.e[i] = {X0.e[i].x * sc, X0.e[i].y * sc};
    Y.e[i] = {Y0.e[i].x * sc, Y0.e[i].y * sc};
  }
  mx_id<D>(E);
#pragma unroll
  for (int i = 0; i < D * D; ++i) F.e[i] = {0.0, 0.0};
  for (int m = TAYLOR - 1; m >= 0; --m) {
    // dP_m = (Y P_{m+1} + X dP_{m+1}) / (m + 1), P_m = I + X P_{m+1} / (m + 1)
    mx_mul<D>(Y, E, T);
    mx_mul<D>(X, F, U);
    const double inv = 1.0 / (m + 1);
#pragma unroll
    for (int i = 0; i < D * D; ++i) F.e[i] = {(T.e[i].x + U.e[i].x) * inv, (T.e[i].y + U.e[i].y) * inv};
    mx_mul<D>(X, E, T);
#pragma unroll
    for (int i = 0; i < D * D; ++i) E.e[i] = {((i / D) == (i % D) ? 1.0 : 0.0) + T.e[i].x * inv, T.e[i].y * inv};
  }
  for (int k = 0; k < s; ++k) {  // (E^2)' = E F + F E
    mx_mul<D>(E, F, T);
    mx_mul<D>(F, E, U);
#pragma unroll
    for (int i = 0; i < D * D; ++i) F.e[i] = caddx(T.e[i], U.e[i]);
    mx_mul<D>(E, E, T);
    E = T;
  }
}

template <int D>
__device__ __forceinline__ void lds_store(cx* L, int slot, const Mx<D>& m) {
#pragma unroll
  for (int i = 0; i < D * D; ++i) L[slot * D * D + i] = m.e[i];
}
template <int D>
__device__ __forceinline__ void lds_load(const cx* L, int slot, Mx<D>& m) {
#pragma unroll
  for (int i = 0; i < D * D; ++i) m.e[i] = L[slot * D * D + i];
}

// lane l: product of its factors, P_l = E_{e-1} ... E_b
template <int D>
__device__ __forceinline__ void chunk_product(const EvolveArgs& A, int g, int b, int e, const cx* H, Mx<D>& P) {
  mx_id<D>(P);
  Mx<D> X, E, T;
  const double* a = A.coef + (int64_t)g * A.n_factors * A.n_terms;
  for (int m = b; m < e; ++m) {
    make_x<D>(a + (int64_t)m * A.n_terms, H, A.n_terms, X);
    expm<D>(X, E);
    mx_mul<D>(E, P, T);
    P = T;
  }
}

constexpr int WAVE = 64;

template <int D>
__global__ __launch_bounds__(WAVE) void evolve_fwd_kernel(EvolveArgs A) {
  __shared__ cx lds[WAVE * D * D];
  const int g = blockIdx.x;
  const int l = threadIdx.x;
  const int M = A.n_factors;
  const int b = (int)((int64_t)M * l / WAVE), e = (int)((int64_t)M * (l + 1) / WAVE);
  const cx* H = reinterpret_cast<const cx*>(A.hmats) + (A.hmats_batched ? (int64_t)g * A.n_terms * D * D : 0);
  Mx<D> P;
  chunk_product<D>(A, g, b, e, H, P);
  lds_store<D>(lds, l, P);
  __syncthreads();
  if (l == 0) {  // U = P_63 ... P_0
    Mx<D> U, Q, T;
    lds_load<D>(lds, 0, U);
    for (int j = 1; j < WAVE; ++j) {
      lds_load<D>(lds, j, Q);
      mx_mul<D>(Q, U, T);
      U = T;
    }
    cx* out = reinterpret_cast<cx*>(A.unitaries) + (int64_t)g * D * D;
#pragma unroll
    for (int i = 0; i < D * D; ++i) out[i] = U.e[i];
  }
}

template <int D>
__global__ __launch_bounds__(WAVE) void evolve_vjp_kernel(EvolveArgs A) {
  __shared__ cx lds[WAVE * D * D];
  const int g = blockIdx.x;
  const int l = threadIdx.x;
  const int M = A.n_factors, K = A.n_terms;
  const int b = (int)((int64_t)M * l / WAVE), e = (int)((int64_t)M * (l + 1) / WAVE);
  const cx* H = reinterpret_cast<const cx*>(A.hmats) + (A.hmats_batched ? (int64_t)g * K * D * D : 0);
  Mx<D> P;
  chunk_product<D>(A, g, b, e, H, P);
  lds_store<D>(lds, l, P);
  __syncthreads();
  // exclusive prefix Pre = P_{l-1} ... P_0 and suffix Suf = P_63 ... P_{l+1}
  Mx<D> Pre, Suf, Q, T;
  mx_id<D>(Pre);
  for (int j = 0; j < l; ++j) {
    lds_load<D>(lds, j, Q);
    mx_mul<D>(Q, Pre, T);
    Pre = T;
  }
  mx_id<D>(Suf);
  for (int j = WAVE - 1; j > l; --j) {
    lds_load<D>(lds, j, Q);
    mx_mul<D>(Suf, Q, T);
    Suf = T;
  }
  // W = L^dag G for the chunk's last factor (L = Suf), R = the product up to
  // and including the chunk (P Pre)
  Mx<D> G, W, R;
  const cx* gu = reinterpret_cast<const cx*>(A.cot) + (int64_t)g * D * D;
#pragma unroll
  for (int i = 0; i < D * D; ++i) G.e[i] = gu[i];
  mx_adj<D>(Suf, T);
  mx_mul<D>(T, G, W);
  mx_mul<D>(P, Pre, R);
  const double* a = A.coef + (int64_t)g * M * K;
  double* ga = A.grad_coef + (int64_t)g * M * K;
  Mx<D> X, E, F, Mm, Y, Ea;
  for (int m = e - 1; m >= b; --m) {
    make_x<D>(a + (int64_t)m * K, H, K, X);
    expm<D>(X, E);
    mx_adj<D>(E, Ea);
    mx_mul<D>(Ea, R, T);  // R_m = E_m^dag R_{m+1}
    R = T;
    mx_adj<D>(R, T);      // M_m = W_m R_m^dag
    mx_mul<D>(W, T, Mm);
    mx_adj<D>(Mm, Y);     // Z = dE[M^dag]
    expm_frechet<D>(X, Y, E, F);
    for (int k = 0; k < K; ++k) {  // Im Tr(H_k Z)
      double s = 0;
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int q = 0; q < D; ++q) {
          const cx h = H[k * D * D + r * D + q], z = F.e[q * D + r];
          s += h.x * z.y + h.y * z.x;
        }
      ga[(int64_t)m * K + k] = s;
    }
    mx_mul<D>(Ea, W, T);  // W_{m-1} = E_m^dag W_m
    W = T;
  }
}

}  // namespace

hipError_t launch_evolve(const EvolveArgs& a, int vjp, hipStream_t st) {
  if (a.n_gates <= 0) return hipSuccess;
  const dim3 grid((unsigned)a.n_gates), block(WAVE);
  if (a.dim == 2) {
    if (vjp)
      hipLaunchKernelGGL(evolve_vjp_kernel<2>, grid, block, 0, st, a);
    else
      hipLaunchKernelGGL(evolve_fwd_kernel<2>, grid, block, 0, st, a);
  } else if (a.dim == 4) {
    if (vjp)
      hipLaunchKernelGGL(evolve_vjp_kernel<4>, grid, block, 0, st, a);
    else
      hipLaunchKernelGGL(evolve_fwd_kernel<4>, grid, block, 0, st, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace jq

// Batched two-qubit pulse propagators (FP64) and their adjoint gradient.
//
// Reference path: Magnus-2 midpoint steps U <- expm(-i h H(t_n + h/2)) U
// (evolution.py:230-235) over n_steps piecewise-constant steps, then the
// process infidelity 1 - |Tr(V^dag U)|^2 / 16 (qoc.py:335-341).  H is
//
//   H(t) = w0/2 Z0 + w1/2 Z1 + J Z0 Z1
//          + E(t) (cos phi X0 + sin phi Y0) + Q(t) (-sin phi X0 + cos phi Y0)
//
// with E the lifted Gaussian (pulses.py:337-352) and Q = -beta dE/dt the DRAG
// quadrature (pulses.py:377-390).  H commutes with Z1, so each step is, per
// value z1 = +-1 of qubit 1, the scalar phase exp(-i h w1 z1 / 2) times
//   exp(-i h (a X + b Y + c_z1 Z)) = cos(h r) I - i sin(h r)/r (a X + b Y + c Z),
//   a = E cos phi - Q sin phi, b = E sin phi + Q cos phi, c_z1 = w0/2 + J z1,
// evaluated in closed form: no matrix exponential is needed.
//
// Parallel layout: one workgroup per parameter set; thread j owns K
// consecutive steps.  Each thread multiplies its steps, the workgroup forms
// inclusive prefix and suffix products in LDS (Hillis-Steele, log2(NT)
// rounds), and the gradient walk goes back over each thread's steps with
//   M_n = R_n V^dag L_n,  d Tr(V^dag U) = sum_n Tr(M_n dU_n),
// R_n = U_{n-1}..U_0 (recovered as U_n^dag R_{n+1}), L_n = U_{N-1}..U_{n+1}.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pulse.h"

namespace jq {

namespace {

constexpr int PNT = 128;  // threads per parameter set

struct cx {
  double r, i;
};
__device__ __forceinline__ cx cmulx(cx a, cx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cx caddx(cx a, cx b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cx conjx(cx a) { return {a.r, -a.i}; }

struct M2 {
  cx m[4];  // row-major 2x2
};

__device__ __forceinline__ M2 ident() {
  M2 o;
  o.m[0] = {1, 0};
  o.m[1] = {0, 0};
  o.m[2] = {0, 0};
  o.m[3] = {1, 0};
  return o;
}

__device__ __forceinline__ M2 mm(const M2& x, const M2& y) {
  M2 o;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) o.m[2 * r + c] = caddx(cmulx(x.m[2 * r], y.m[c]), cmulx(x.m[2 * r + 1], y.m[2 + c]));
  return o;
}

// x^dagger y
__device__ __forceinline__ M2 mhm(const M2& x, const M2& y) {
  M2 o;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
      o.m[2 * r + c] = caddx(cmulx(conjx(x.m[r]), y.m[c]), cmulx(conjx(x.m[2 + r]), y.m[2 + c]));
  return o;
}

// Tr(x y)
__device__ __forceinline__ cx trmm(const M2& x, const M2& y) {
  cx t = {0, 0};
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int k = 0; k < 2; ++k) t = caddx(t, cmulx(x.m[2 * r + k], y.m[2 * k + r]));
  return t;
}

struct Env {
  double a, b;            // drive coefficients of X0, Y0
  double da[4], db[4];    // d/d(A, sigma, beta, phi)
};

// Envelopes at time t and their parameter derivatives.
__device__ __forceinline__ Env envelope(double t, double A, double sig, double beta, double cphi, double sphi,
                                        double tc, bool grad) {
  const double s2 = sig * sig;
  const double x = t - tc;
  const double aa = -0.5 * x * x / s2;          // log g(t)
  const double bb = -0.5 * tc * tc / s2;        // log g(0)
  const double d = t * (2.0 * tc - t) / (2.0 * s2);
  const double rise = expm1(fmin(d, 0.0)) - expm1(fmin(-d, 0.0));
  const double lift = -expm1(bb);               // 1 - g(0)
  const double lg = exp(fmax(aa, bb)) * rise / lift;
  const double g = exp(aa);
  const double q1 = x / s2 * g / lift;          // Q / (A beta)
  const double E = A * lg, Q = A * beta * q1;
  Env e;
  e.a = E * cphi - Q * sphi;
  e.b = E * sphi + Q * cphi;
  if (grad) {
    // A: E and Q are linear in A
    e.da[0] = lg * cphi - beta * q1 * sphi;
    e.db[0] = lg * sphi + beta * q1 * cphi;
    // sigma
    const double g0 = exp(bb);
    const double dg = g * x * x / (s2 * sig);
    const double dg0 = g0 * tc * tc / (s2 * sig);
    const double dlg = ((dg - dg0) * lift + (g - g0) * dg0) / (lift * lift);
    const double dq1 = x * (dg / s2 - 2.0 * g / (s2 * sig) + g / s2 * dg0 / lift) / lift;
    const double dE = A * dlg, dQ = A * beta * dq1;
    e.da[1] = dE * cphi - dQ * sphi;
    e.db[1] = dE * sphi + dQ * cphi;
    // beta
    e.da[2] = -A * q1 * sphi;
    e.db[2] = A * q1 * cphi;
    // phi
    e.da[3] = -e.b;
    e.db[3] = e.a;
  }
  return e;
}

// Phase-free step exp(-i h (a X + b Y + c Z)) and, optionally, its a / b derivatives.
__device__ __forceinline__ void step_mat(double h, double a, double b, double c, M2& U, M2* dUa, M2* dUb) {
  const double r = sqrt(a * a + b * b + c * c);
  double S, Cc;
  sincos(h * r, &S, &Cc);
  const double v = r > 0 ? S / r : h;
  U.m[0] = {Cc, -v * c};
  U.m[1] = {-v * b, -v * a};
  U.m[2] = {v * b, -v * a};
  U.m[3] = {Cc, v * c};
  if (dUa) {
    // du/dp = -h S p / r ; dv/dp = p (h r C - S) / r^3   (p = a or b)
    const double r3 = r * r * r;
    const double kv = r > 0 ? (h * r * Cc - S) / r3 : 0.0;
    const double ku = r > 0 ? -h * S / r : 0.0;
    // dU = du I - i (dv K + v G),  K = [[c, a - ib], [a + ib, -c]]
    {
      const double du = ku * a, dv = kv * a;  // G = X
      dUa->m[0] = {du, -dv * c};
      dUa->m[1] = {-dv * b, -(dv * a + v)};
      dUa->m[2] = {dv * b, -(dv * a + v)};
      dUa->m[3] = {du, dv * c};
    }
    {
      const double du = ku * b, dv = kv * b;  // G = Y = [[0, -i], [i, 0]]: -i v Y = [[0, -v], [v, 0]]
      dUb->m[0] = {du, -dv * c};
      dUb->m[1] = {-dv * b - v, -dv * a};
      dUb->m[2] = {dv * b + v, -dv * a};
      dUb->m[3] = {du, dv * c};
    }
  }
}

__global__ __launch_bounds__(PNT) void pulse_kernel(PulseArgs p) {
  __shared__ M2 pre[2][PNT];
  __shared__ M2 suf[2][PNT];
  __shared__ double red[8][PNT / 64];
  const int set = blockIdx.x;
  const int j = threadIdx.x;
  const double* pp = p.params + (int64_t)set * 6;
  const double A = pp[0], sig = pp[1], beta = pp[2], phi = pp[3], w0 = pp[4], w1 = pp[5];
  double cphi, sphi;
  sincos(phi, &sphi, &cphi);
  const int N = p.n_steps;
  const double h = p.T / N;
  const double tc = 0.5 * p.T;
  const int K = (N + PNT - 1) / PNT;
  const int n0 = j * K, n1 = min(N, n0 + K);
  const double cz[2] = {0.5 * w0 + p.J, 0.5 * w0 - p.J};  // q1 = 0 (z1 = +1), q1 = 1 (z1 = -1)
  // 1. local products (later steps on the left)
  M2 loc[2] = {ident(), ident()};
  for (int n = n0; n < n1; ++n) {
    const Env e = envelope((n + 0.5) * h, A, sig, beta, cphi, sphi, tc, false);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      M2 U;
      step_mat(h, e.a, e.b, cz[s], U, nullptr, nullptr);
      loc[s] = mm(U, loc[s]);
    }
  }
  // 2. inclusive prefix (U_j..U_0 block products) and suffix (U_{NT-1}..U_j)
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    pre[s][j] = loc[s];
    suf[s][j] = loc[s];
  }
  __syncthreads();
  for (int d = 1; d < PNT; d <<= 1) {
    M2 x[2], y[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      x[s] = j >= d ? mm(pre[s][j], pre[s][j - d]) : pre[s][j];
      y[s] = j + d < PNT ? mm(suf[s][j + d], suf[s][j]) : suf[s][j];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      pre[s][j] = x[s];
      suf[s][j] = y[s];
    }
    __syncthreads();
  }
  // total propagator, per block, phase-free; the block phases exp(-i T w1 z1 / 2)
  M2 Ut[2] = {suf[0][0], suf[1][0]};
  double ph_r[2], ph_i[2];
  sincos(-0.5 * p.T * w1, &ph_i[0], &ph_r[0]);
  ph_r[1] = ph_r[0];
  ph_i[1] = -ph_i[0];
  // target blocks V_s = V restricted to q1 = s (indices 2 q0 + s)
  M2 Vs[2];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c) Vs[s].m[2 * r + c] = {p.V[2 * ((2 * r + s) * 4 + 2 * c + s)], p.V[2 * ((2 * r + s) * 4 + 2 * c + s) + 1]};
  cx tr = {0, 0};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const cx ts = trmm(mhm(Vs[s], Ut[s]), ident());  // Tr(V_s^dag U_s)
    tr = caddx(tr, cmulx({ph_r[s], ph_i[s]}, ts));
  }
  if (j == 0) {
    p.infid[set] = 1.0 - (tr.r * tr.r + tr.i * tr.i) / 16.0;
    if (p.U) {
      double* Uo = p.U + (int64_t)set * 32;
      for (int e = 0; e < 32; ++e) Uo[e] = 0.0;
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            const cx u = cmulx({ph_r[s], ph_i[s]}, Ut[s].m[2 * r + c]);
            const int e = (2 * r + s) * 4 + 2 * c + s;
            Uo[2 * e] = u.r;
            Uo[2 * e + 1] = u.i;
          }
    }
  }
  if (!p.grad) return;
  // 3. gradient walk over this thread's steps, last to first
  cx dt[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  M2 R[2], L[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    R[s] = pre[s][j];                              // U_{n1-1} .. U_0
    L[s] = j + 1 < PNT ? suf[s][j + 1] : ident();  // U_{N-1} .. U_{n1}
  }
  for (int n = n1 - 1; n >= n0; --n) {
    const Env e = envelope((n + 0.5) * h, A, sig, beta, cphi, sphi, tc, true);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      M2 U, dA, dB;
      step_mat(h, e.a, e.b, cz[s], U, &dA, &dB);
      R[s] = mhm(U, R[s]);                         // U_{n-1} .. U_0
      const M2 Mn = mm(R[s], mhm(Vs[s], L[s]));    // R_n V^dag L_n
      const cx ga = cmulx({ph_r[s], ph_i[s]}, trmm(Mn, dA));
      const cx gb = cmulx({ph_r[s], ph_i[s]}, trmm(Mn, dB));
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        dt[k].r += ga.r * e.da[k] + gb.r * e.db[k];
        dt[k].i += ga.i * e.da[k] + gb.i * e.db[k];
      }
      L[s] = mm(L[s], U);
    }
  }
  // d infidelity / dp = -(2/16) Re(conj(tr) dTr/dp), reduced over the workgroup
  double g[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) g[k] = -(tr.r * dt[k].r + tr.i * dt[k].i) / 8.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double v = g[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((j & 63) == 0) red[k][j >> 6] = v;
  }
  __syncthreads();
  if (j < 4) {
    double v = 0;
    for (int w = 0; w < PNT / 64; ++w) v += red[j][w];
    p.grad[(int64_t)set * 4 + j] = v;
  }
}

}  // namespace

hipError_t launch_pulse(const PulseArgs& a, hipStream_t st) {
  if (a.n_sets <= 0) return hipSuccess;
  hipLaunchKernelGGL(pulse_kernel, dim3(a.n_sets), dim3(PNT), 0, st, a);
  return hipGetLastError();
}

}  // namespace jq

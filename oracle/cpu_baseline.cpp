// C++/OpenMP restatement of the reference statevector algorithm, used as the
// CPU timing baseline.  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
//
// Follows the reference plan for the hardware-efficient ansatz exactly:
//  - _fuse (jaqsi/simulation.py:214-240): RX, RY, RZ on one wire (and an
//    encoding RX in front) multiply into one 2x2 block per wire and layer;
//  - _fuse_runs (simulation.py:290-335): the CX ring of a layer is a run of
//    permutations, composed into one 2^n int32 gather (and its inverse);
//  - _run (simulation.py:400-409): one pass over the state per plan entry;
//  - the adjoint of simulation.py:616-652: psi <- U^dagger psi,
//    bar[o][i] = sum lambda[o] psi[i], lambda <- U^T lambda, permutations by
//    their inverse gather; bars chained to the angles.
// Arithmetic is complex64 (the reference's default dtype) or complex128 (its
// jax_enable_x64 mode, operations.py:26-43), reductions in FP64.
#include <cmath>
#include <algorithm>
#include <array>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

typedef std::complex<double> cdd;
typedef std::complex<float> cf32;

namespace {

struct M2 {
  cdd a[4];
};

M2 mul(const M2& x, const M2& y) {
  M2 r;
  r.a[0] = x.a[0] * y.a[0] + x.a[1] * y.a[2];
  r.a[1] = x.a[0] * y.a[1] + x.a[1] * y.a[3];
  r.a[2] = x.a[2] * y.a[0] + x.a[3] * y.a[2];
  r.a[3] = x.a[2] * y.a[1] + x.a[3] * y.a[3];
  return r;
}

// cos(t/2) I - i sin(t/2) P and its derivative (gateset.py:399-405)
M2 rot(int axis, double t, bool deriv) {
  double c = std::cos(t / 2), s = std::sin(t / 2);
  if (deriv) {
    double c2 = -0.5 * s, s2 = 0.5 * c;
    c = c2;
    s = s2;
  }
  const cdd I(0, 1);
  M2 m;
  if (axis == 0) {  // X
    m.a[0] = c; m.a[1] = -I * s; m.a[2] = -I * s; m.a[3] = c;
  } else if (axis == 1) {  // Y
    m.a[0] = c; m.a[1] = -s; m.a[2] = s; m.a[3] = c;
  } else {  // Z
    m.a[0] = cdd(c, -s); m.a[1] = 0; m.a[2] = 0; m.a[3] = cdd(c, s);
  }
  return m;
}

template <typename cf>
void apply2(cf* psi, int64_t dim, int bit, const M2& m) {
  const cf m0(m.a[0]), m1(m.a[1]), m2(m.a[2]), m3(m.a[3]);
  const int64_t stride = 1LL << bit;
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < dim / 2; ++g) {
    const int64_t i = ((g >> bit) << (bit + 1)) | (g & (stride - 1));
    const int64_t j = i | stride;
    const cf x = psi[i], y = psi[j];
    psi[i] = m0 * x + m1 * y;
    psi[j] = m2 * x + m3 * y;
  }
}

template <typename cf>
void gather(cf* psi, cf* tmp, const int32_t* idx, int64_t dim) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) tmp[i] = psi[idx[i]];
  std::memcpy(psi, tmp, dim * sizeof(cf));
}

// Forward + adjoint gradient of sum_q <Z_q> for the hardware-efficient
// ansatz: theta[layers][n][3] (RX, RY, RZ), optional features[n] (encoding RX).
// Returns the expectation values (n) and d/dtheta (layers*n*3), in FP64.
template <typename cf>
int hea_grad(int n, int layers, const double* theta, const double* features, int threads, double* expvals,
             double* grads) {
  typedef typename cf::value_type rf;
  if (threads > 0) omp_set_num_threads(threads);
  const int64_t dim = 1LL << n;
  std::vector<cf> psi(dim), lam(dim), tmp(dim);
  // composed CX ring gather: out[i] = in[g_1(g_2(...g_n(i)))]
  std::vector<int32_t> idx(dim), inv(dim);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) {
    int64_t x = i;
    for (int q = n - 1; q >= 0; --q) {
      const int c = n - 1 - q, t = n - 1 - ((q + 1) % n);
      if ((x >> c) & 1) x ^= 1LL << t;
    }
    idx[i] = (int32_t)x;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) inv[idx[i]] = (int32_t)i;
  // fused blocks per (layer, wire)
  std::vector<M2> blocks(layers * n);
  for (int l = 0; l < layers; ++l)
    for (int q = 0; q < n; ++q) {
      const double* t = theta + (l * n + q) * 3;
      M2 m = mul(rot(2, t[2], false), mul(rot(1, t[1], false), rot(0, t[0], false)));
      if (l == 0 && features) m = mul(m, rot(0, features[q], false));
      blocks[l * n + q] = m;
    }
  // forward
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) psi[i] = 0;
  psi[0] = 1;
  for (int l = 0; l < layers; ++l) {
    for (int q = 0; q < n; ++q) apply2(psi.data(), dim, n - 1 - q, blocks[l * n + q]);
    gather(psi.data(), tmp.data(), idx.data(), dim);
  }
  // expectation values and the lambda seed 2 conj(psi) sum_q z_q
  std::vector<double> ev(n, 0.0);
#pragma omp parallel
  {
    std::vector<double> loc(n, 0.0);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < dim; ++i) {
      const double p = std::norm(psi[i]);
      double w = 0;
      for (int q = 0; q < n; ++q) {
        const double z = ((i >> (n - 1 - q)) & 1) ? -1.0 : 1.0;
        loc[q] += p * z;
        w += z;
      }
      lam[i] = cf(rf(2) * (rf)w) * std::conj(psi[i]);
    }
#pragma omp critical
    for (int q = 0; q < n; ++q) ev[q] += loc[q];
  }
  for (int q = 0; q < n; ++q) expvals[q] = ev[q];
  // backward over the reversed plan
  std::vector<M2> bars(layers * n);
  for (int l = layers - 1; l >= 0; --l) {
    gather(psi.data(), tmp.data(), inv.data(), dim);
    gather(lam.data(), tmp.data(), inv.data(), dim);
    for (int q = n - 1; q >= 0; --q) {
      const M2& m = blocks[l * n + q];
      M2 adj, tr;
      adj.a[0] = std::conj(m.a[0]); adj.a[1] = std::conj(m.a[2]);
      adj.a[2] = std::conj(m.a[1]); adj.a[3] = std::conj(m.a[3]);
      tr.a[0] = m.a[0]; tr.a[1] = m.a[2]; tr.a[2] = m.a[1]; tr.a[3] = m.a[3];
      const int bit = n - 1 - q;
      apply2(psi.data(), dim, bit, adj);
      const int64_t stride = 1LL << bit;
      double b[8] = {0};
#pragma omp parallel
      {
        double lb[8] = {0};
#pragma omp for schedule(static)
        for (int64_t g = 0; g < dim / 2; ++g) {
          const int64_t i = ((g >> bit) << (bit + 1)) | (g & (stride - 1));
          const int64_t j = i | stride;
          const cf l0 = lam[i], l1 = lam[j], p0 = psi[i], p1 = psi[j];
          const cf v[4] = {l0 * p0, l0 * p1, l1 * p0, l1 * p1};
          for (int e = 0; e < 4; ++e) {
            lb[2 * e] += v[e].real();
            lb[2 * e + 1] += v[e].imag();
          }
        }
#pragma omp critical
        for (int e = 0; e < 8; ++e) b[e] += lb[e];
      }
      for (int e = 0; e < 4; ++e) bars[l * n + q].a[e] = cdd(b[2 * e], b[2 * e + 1]);
      apply2(lam.data(), dim, bit, tr);
    }
  }
  // chain rule through the fused products (JAX's transpose of _rot_matrix)
  for (int l = 0; l < layers; ++l)
    for (int q = 0; q < n; ++q) {
      const double* t = theta + (l * n + q) * 3;
      M2 f[4];
      int nf = 0;
      if (l == 0 && features) f[nf++] = rot(0, features[q], false);
      const int first = nf;
      for (int a = 0; a < 3; ++a) f[nf++] = rot(a, t[a], false);
      for (int a = 0; a < 3; ++a) {
        const int j = first + a;
        M2 L = {{1, 0, 0, 1}}, R = {{1, 0, 0, 1}};
        for (int k = j + 1; k < nf; ++k) L = mul(f[k], L);
        for (int k = 0; k < j; ++k) R = mul(f[k], R);
        const M2 d = mul(L, mul(rot(a, t[a], true), R));
        double g = 0;
        for (int e = 0; e < 4; ++e) g += (bars[l * n + q].a[e] * d.a[e]).real();
        grads[(l * n + q) * 3 + a] = g;
      }
    }
  return 0;
}

// ---------------------------------------------------------------------------
// Density matrix (the reference's _compile_mixed plan, simulation.py:348-381,
// 443-473): rho as a 2n-wire complex64 vector (ket wire w = bit 2n-1-w, bra
// wire w = bit n-1-w); per layer and wire one fused superoperator block
// RX.dep.amp.RY.dep.amp.RZ.dep.amp (_fuse merges every entry on the wire
// pair (w, w+n)); each CX of the ring as one gather over ket and bra bits
// (a run of two permutations, _fuse_runs), then the fused dep.amp blocks of
// its two wires.
// ---------------------------------------------------------------------------
struct M4 {
  cdd a[16];
};

M4 mul4(const M4& x, const M4& y) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cdd s = 0;
      for (int k = 0; k < 4; ++k) s += x.a[4 * i + k] * y.a[4 * k + j];
      r.a[4 * i + j] = s;
    }
  return r;
}

M4 superop_unitary(const M2& u) {  // kron(U, conj U)
  M4 r;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k)
        for (int l = 0; l < 2; ++l) r.a[4 * (2 * i + k) + (2 * j + l)] = u.a[2 * i + j] * std::conj(u.a[2 * k + l]);
  return r;
}

M4 superop_kraus(const std::vector<M2>& ks) {  // sum kron(K, conj K)
  M4 r;
  for (auto& x : r.a) x = 0;
  for (const M2& k : ks) {
    M4 t = superop_unitary(k);
    for (int e = 0; e < 16; ++e) r.a[e] += t.a[e];
  }
  return r;
}

M4 dep_amp(double p, double g) {  // AmplitudeDamping(g) after DepolarizingChannel(p)
  const cdd I(0, 1);
  const double s0 = std::sqrt(1 - p), s1 = std::sqrt(p / 3);
  std::vector<M2> dep = {{{s0, 0, 0, s0}}, {{0, s1, s1, 0}}, {{0, -I * s1, I * s1, 0}}, {{s1, 0, 0, -s1}}};
  std::vector<M2> amp = {{{1, 0, 0, std::sqrt(1 - g)}}, {{0, std::sqrt(g), 0, 0}}};
  return mul4(superop_kraus(amp), superop_kraus(dep));
}

void apply4(cf32* v, int64_t dim, int bit_hi, int bit_lo, const M4& m) {
  std::complex<float> mf[16];
  for (int e = 0; e < 16; ++e) mf[e] = std::complex<float>(m.a[e]);
  const int64_t shi = 1LL << bit_hi, slo = 1LL << bit_lo;
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < dim / 4; ++g) {
    // insert zeros at bit_lo then bit_hi (bit_hi > bit_lo)
    int64_t i = g;
    i = ((i >> bit_lo) << (bit_lo + 1)) | (i & (slo - 1));
    i = ((i >> bit_hi) << (bit_hi + 1)) | (i & (shi - 1));
    const int64_t idx[4] = {i, i | slo, i | shi, i | shi | slo};
    std::complex<float> x[4], y[4];
    for (int r = 0; r < 4; ++r) x[r] = v[idx[r]];
    for (int r = 0; r < 4; ++r) {
      std::complex<float> s = 0;
      for (int c = 0; c < 4; ++c) s += mf[4 * r + c] * x[c];
      y[r] = s;
    }
    for (int r = 0; r < 4; ++r) v[idx[r]] = y[r];
  }
}

// ---------------------------------------------------------------------------
// Pulse propagators: the reference's Magnus-2 solve (evolution.py:230-235,
// one expm of the 4x4 step generator per step) of the two-qubit DRAG drive,
// the process infidelity 1 - |Tr(V^dag U)|^2 / 16 (qoc.py:335-341), and its
// gradient w.r.t. (A, sigma, beta, phi) -- what reverse-mode autodiff of the
// scan computes -- by the adjoint sweep over the steps with the Frechet
// derivative of each step's exponential.  FP64.
// ---------------------------------------------------------------------------
typedef std::array<cdd, 16> Q4;

Q4 q_mul(const Q4& x, const Q4& y) {
  Q4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cdd s = 0;
      for (int k = 0; k < 4; ++k) s += x[4 * i + k] * y[4 * k + j];
      r[4 * i + j] = s;
    }
  return r;
}
Q4 q_id() {
  Q4 r;
  for (int e = 0; e < 16; ++e) r[e] = (e % 5 == 0) ? 1.0 : 0.0;
  return r;
}
Q4 q_adj(const Q4& x) {
  Q4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[4 * i + j] = std::conj(x[4 * j + i]);
  return r;
}
int q_scaling(const Q4& X) {
  double n1 = 0;
  for (int j = 0; j < 4; ++j) {
    double c = 0;
    for (int i = 0; i < 4; ++i) c += std::abs(X[4 * i + j]);
    n1 = std::max(n1, c);
  }
  int s = 0;
  while (n1 > 0.25 && s < 60) n1 *= 0.5, ++s;
  return s;
}
// exp(X) (Taylor degree 16 after scaling by 2^-s, s squarings); with F the
// Frechet derivative in direction Y when wanted
void q_expm(const Q4& X0, const Q4* Y0, Q4& E, Q4* F) {
  const int s = q_scaling(X0);
  const double sc = std::ldexp(1.0, -s);
  Q4 X, Y;
  for (int e = 0; e < 16; ++e) X[e] = X0[e] * sc;
  if (F) for (int e = 0; e < 16; ++e) Y[e] = (*Y0)[e] * sc;
  E = q_id();
  Q4 D;
  for (auto& x : D) x = 0;
  for (int m = 15; m >= 0; --m) {
    const double inv = 1.0 / (m + 1);
    if (F) {
      Q4 a = q_mul(Y, E), b = q_mul(X, D);
      for (int e = 0; e < 16; ++e) D[e] = (a[e] + b[e]) * inv;
    }
    Q4 t = q_mul(X, E);
    for (int e = 0; e < 16; ++e) E[e] = ((e % 5 == 0) ? 1.0 : 0.0) + t[e] * inv;
  }
  for (int k = 0; k < s; ++k) {
    if (F) {
      Q4 a = q_mul(E, D), b = q_mul(D, E);
      for (int e = 0; e < 16; ++e) D[e] = a[e] + b[e];
    }
    E = q_mul(E, E);
  }
  if (F) *F = D;
}

struct PulseSet {
  double A, sigma, beta, phi, w0, w1;
};

void drag_env(const PulseSet& p, double sigma, double t, double tc, double& E, double& Q) {
  const double a = -0.5 * ((t - tc) / sigma) * ((t - tc) / sigma);
  const double b = -0.5 * (tc / sigma) * (tc / sigma);
  const double d = t * (2 * tc - t) / (2 * sigma * sigma);
  const double rise = std::expm1(std::min(d, 0.0)) - std::expm1(std::min(-d, 0.0));
  const double lift = -std::expm1(b);
  E = p.A * std::exp(std::max(a, b)) * rise / lift;
  Q = p.A * p.beta * (t - tc) / (sigma * sigma) * std::exp(a) / lift;
}

// generator X = -i h H(t) of one step; also the (a, b) drive coefficients
Q4 step_generator(const PulseSet& p, double J, double h, double t, double T, double& a, double& b) {
  double E, Q;
  drag_env(p, p.sigma, t, T / 2, E, Q);
  const double c = std::cos(p.phi), sn = std::sin(p.phi);
  a = E * c - Q * sn;
  b = E * sn + Q * c;
  // H = w0/2 Z0 + w1/2 Z1 + J Z0 Z1 + a X0 + b Y0 (wire 0 = most significant)
  Q4 H;
  for (auto& x : H) x = 0;
  const double z[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
  for (int i = 0; i < 4; ++i) H[5 * i] = 0.5 * p.w0 * z[i][0] + 0.5 * p.w1 * z[i][1] + J * z[i][0] * z[i][1];
  const cdd I(0, 1);
  for (int q1 = 0; q1 < 2; ++q1) {  // X0, Y0 couple rows q1 and 2 + q1
    H[4 * q1 + 2 + q1] += a - I * b;
    H[4 * (2 + q1) + q1] += a + I * b;
  }
  Q4 X;
  for (int e = 0; e < 16; ++e) X[e] = -I * h * H[e];
  return X;
}

}  // namespace

extern "C" {

// params[n_sets][6] = (A, sigma, beta, phi, w0, w1); target: 4x4 complex128
// row-major, re/im interleaved.  infid[n_sets], grad[n_sets][4].
int cpu_pulse_grad(int n_sets, const double* params, int n_steps, double T, double J, const double* target,
                   int threads, double* infid, double* grad) {
  if (threads > 0) omp_set_num_threads(threads);
  Q4 V;
  for (int e = 0; e < 16; ++e) V[e] = cdd(target[2 * e], target[2 * e + 1]);
  const Q4 Vd = q_adj(V);
  const double h = T / n_steps;
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < n_sets; ++s) {
    const double* q = params + 6 * s;
    const PulseSet p{q[0], q[1], q[2], q[3], q[4], q[5]};
    Q4 U = q_id(), E;
    double a, b;
    for (int n = 0; n < n_steps; ++n) {
      q_expm(step_generator(p, J, h, (n + 0.5) * h, T, a, b), nullptr, E, nullptr);
      U = q_mul(E, U);
    }
    cdd tau = 0;
    const Q4 VU = q_mul(Vd, U);
    for (int i = 0; i < 4; ++i) tau += VU[5 * i];
    infid[s] = 1.0 - std::norm(tau) / 16.0;
    // dL = Re Tr(G^dag dU) with G = -(tau / 8) V
    Q4 W, R = U;
    for (int e = 0; e < 16; ++e) W[e] = -(tau / 8.0) * V[e];
    double g[4] = {0, 0, 0, 0};
    const double ds = 1e-6 * p.sigma;
    for (int n = n_steps - 1; n >= 0; --n) {
      const double t = (n + 0.5) * h;
      const Q4 X = step_generator(p, J, h, t, T, a, b);
      q_expm(X, nullptr, E, nullptr);
      const Q4 Ea = q_adj(E);
      R = q_mul(Ea, R);                        // R_n
      const Q4 M = q_mul(W, q_adj(R));         // M_n = L_n^dag G R_n^dag
      const Q4 Md = q_adj(M);
      Q4 Ed, Z;
      q_expm(X, &Md, Ed, &Z);                  // Z = dE[M^dag]
      // dL/da = Re Tr(-i h X0 Z) = h Im Tr(X0 Z), likewise Y0
      cdd tx = 0, ty = 0;
      const cdd I(0, 1);
      for (int q1 = 0; q1 < 2; ++q1) {
        tx += Z[4 * (2 + q1) + q1] + Z[4 * q1 + 2 + q1];
        ty += -I * Z[4 * (2 + q1) + q1] + I * Z[4 * q1 + 2 + q1];  // (Y0)_{q1,2+q1} = -i, (Y0)_{2+q1,q1} = i
      }
      const double la = h * tx.imag(), lb = h * ty.imag();
      // chain rule to (A, sigma, beta, phi); the sigma derivative of the
      // envelopes by a central difference of step 1e-6 sigma
      double E0, Q0, Ep, Qp, Em, Qm;
      drag_env(p, p.sigma, t, T / 2, E0, Q0);
      drag_env(p, p.sigma + ds, t, T / 2, Ep, Qp);
      drag_env(p, p.sigma - ds, t, T / 2, Em, Qm);
      const double c = std::cos(p.phi), sn = std::sin(p.phi);
      auto ab = [&](double dE, double dQ, double& da, double& db) {
        da = dE * c - dQ * sn;
        db = dE * sn + dQ * c;
      };
      double da, db;
      ab(E0 / p.A, Q0 / p.A, da, db);
      g[0] += la * da + lb * db;
      ab((Ep - Em) / (2 * ds), (Qp - Qm) / (2 * ds), da, db);
      g[1] += la * da + lb * db;
      ab(0.0, p.beta != 0.0 ? Q0 / p.beta : 0.0, da, db);
      g[2] += la * da + lb * db;
      g[3] += la * (-b) + lb * a;
      W = q_mul(Ea, W);                        // W_{n-1} = E_n^dag W_n
    }
    for (int k = 0; k < 4; ++k) grad[4 * s + k] = g[k];
  }
  return 0;
}

// Forward of the noisy density-matrix ansatz: Tr(Z_0 rho) after `layers`
// layers, DepolarizingChannel(p_dep) then AmplitudeDamping(gamma) after every
// gate on each touched wire.  theta[layers][n][3].
int cpu_dm_forward(int n, int layers, const double* theta, double p_dep, double gamma, int threads, double* z0) {
  if (threads > 0) omp_set_num_threads(threads);
  const int nn = 2 * n;
  const int64_t dim = 1LL << nn;
  std::vector<cf32> rho(dim), tmp(dim);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) rho[i] = 0;
  rho[0] = 1;
  const M4 da = dep_amp(p_dep, gamma);
  for (int l = 0; l < layers; ++l) {
    for (int q = 0; q < n; ++q) {
      const double* t = theta + (l * n + q) * 3;
      M4 blk = superop_unitary(rot(0, t[0], false));
      blk = mul4(da, blk);
      blk = mul4(superop_unitary(rot(1, t[1], false)), blk);
      blk = mul4(da, blk);
      blk = mul4(superop_unitary(rot(2, t[2], false)), blk);
      blk = mul4(da, blk);
      apply4(rho.data(), dim, nn - 1 - q, n - 1 - q, blk);
    }
    for (int q = 0; q < n; ++q) {
      const int c = q, t = (q + 1) % n;
      const int kc = nn - 1 - c, kt = nn - 1 - t, bc = n - 1 - c, bt = n - 1 - t;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < dim; ++i) {
        int64_t x = i;
        if ((x >> kc) & 1) x ^= 1LL << kt;
        if ((x >> bc) & 1) x ^= 1LL << bt;
        tmp[i] = rho[x];
      }
      std::memcpy(rho.data(), tmp.data(), dim * sizeof(cf32));
      apply4(rho.data(), dim, nn - 1 - c, n - 1 - c, da);
      apply4(rho.data(), dim, nn - 1 - t, n - 1 - t, da);
    }
  }
  const int64_t d1 = 1LL << n;
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < d1; ++i) s += (((i >> (n - 1)) & 1) ? -1.0 : 1.0) * rho[i * d1 + i].real();
  *z0 = s;
  return 0;
}

int cpu_hea_grad(int n, int layers, const double* theta, const double* features, int threads, double* expvals,
                 double* grads) {
  return hea_grad<std::complex<float>>(n, layers, theta, features, threads, expvals, grads);
}

// The same algorithm in complex128 (the reference's x64 mode).
int cpu_hea_grad_c128(int n, int layers, const double* theta, const double* features, int threads, double* expvals,
                      double* grads) {
  return hea_grad<std::complex<double>>(n, layers, theta, features, threads, expvals, grads);
}

}  // extern "C"

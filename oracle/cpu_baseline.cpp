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
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

typedef std::complex<double> cdd;

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

}  // namespace

extern "C" {

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

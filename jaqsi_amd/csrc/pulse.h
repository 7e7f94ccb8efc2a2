// Launch interface of the batched pulse-propagator kernel (pulse.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jq {

struct PulseArgs {
  int n_sets;
  int n_steps;
  double T;              // total time (ns)
  double J;              // ZZ coupling (rad/ns)
  const double* params;  // [n_sets][6]: A, sigma, beta, phi, w0, w1
  double V[32];          // target 4x4 complex128, row-major, (re, im) interleaved
  double* U;             // [n_sets][4][4] complex128 or nullptr
  double* infid;         // [n_sets]
  double* grad;          // [n_sets][4] d infidelity / d(A, sigma, beta, phi) or nullptr
};

hipError_t launch_pulse(const PulseArgs& a, hipStream_t st);

}  // namespace jq

// Launch interface of the batched propagator kernels (evolve.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jq {

struct EvolveArgs {
  int n_gates;        // G (one wavefront each)
  int dim;            // 2 or 4
  int n_terms;        // K
  int n_factors;      // M
  int hmats_batched;  // H_k per gate ([G][K][d][d]) or shared ([K][d][d])
  const double* coef; // [G][M][K] real coefficients a
  const void* hmats;  // complex128 Hermitian terms
  void* unitaries;    // forward output [G][d][d] complex128
  const void* cot;    // VJP input: dL/dRe U + i dL/dIm U, [G][d][d] complex128
  double* grad_coef;  // VJP output: dL/da, [G][M][K]
};

hipError_t launch_evolve(const EvolveArgs& a, int vjp, hipStream_t st);

}  // namespace jq

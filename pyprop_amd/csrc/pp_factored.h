// Factored Hamiltonian H = sum_t f_t A_t (x) R_t on rank-3 (channel, r1, r2) tensors.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pp_common.h"

namespace pp {

constexpr int kMaxFTerms = 32;

struct FHam {
  int nc = 0, n = 0, nterms = 0, bw = 0, npot = 0, Kp = 0;  // Kp: angular GEMM K = nterms * nc rounded up to 4
  std::vector<int> term_pot;          // potential (time scaling) index of each term
  DeviceBuffer<double> A64, bt64;     // A: [ceil16(nc)][Kp] (row c, column t*nc + c);
  DeviceBuffer<float> A32, bt32;      // bt: band tiles in MFMA lane order [T][n][w][tiles][KS][64]
  mutable DeviceBuffer<double> W64;   // radial products Z [T][nc][n][n] complex
  mutable DeviceBuffer<float> W32;
  mutable DeviceBuffer<double> X64;   // channel-minor copy of the input
  mutable DeviceBuffer<float> X32;
  // dst = addend + alpha * sum_t scalings[term_pot[t]] (A_t (x) R_t) src
  void apply64(const double* scalings, const double2* src, double2* dst, const double2* addend, double2 alpha,
               hipStream_t s) const;
  void apply32(const double* scalings, const float2* src, float2* dst, const float2* addend, double2 alpha,
               hipStream_t s) const;
};

}  // namespace pp

struct pp_fham {
  pp::FHam f;
};

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
  // skip_radial (device int, may be NULL): when it reads non-zero, the radial products of
  // the previous FP64 multiply are reused (src must then hold the same values as that call's)
  void apply64(const double* scalings, const double2* src, double2* dst, const double2* addend, double2 alpha,
               hipStream_t s, const int* skip_radial = nullptr) const;
  mutable unsigned long long gen64 = 0;  // FP64 multiplies issued (invalidates reuse)
  void apply32(const double* scalings, const float2* src, float2* dst, const float2* addend, double2 alpha,
               hipStream_t s) const;
};

// *flag = 1 if a and b differ anywhere (n elements), else 0
void fham_flag_differs(long long n, const double2* a, const double2* b, int* flag, hipStream_t s);

}  // namespace pp

struct pp_fham {
  pp::FHam f;
};

// Factored Hamiltonian H = sum_t f_t A_t (x) R_t on rank-3 (channel, r1, r2) tensors.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pp_common.h"

namespace pp {

constexpr int kMaxFTerms = 32;
constexpr int kFhamRadialVariants = 15;  // radial-stage kernel shapes (band-tile and position forms)
constexpr int kFhamAngularVariants = 7;  // angular-stage forms (GEMM: waves splitting K, row blocks per workgroup; sparse)

struct FHamOut {
  double2* dst = nullptr;
  const double2* addend = nullptr;
  double2 alpha = make_double2(1.0, 0.0), beta = make_double2(0.0, 0.0);
};

struct FHam {
  int nc = 0, n = 0, nterms = 0, bw = 0, npot = 0, Kp = 0;  // Kp: angular GEMM K = nterms * nc rounded up to 4
  std::vector<int> term_pot;          // potential (time scaling) index of each term
  DeviceBuffer<int> tk;               // term index of each angular GEMM column k (k / nc)
  DeviceBuffer<double> A64, bt64;     // A: MFMA operand order [ceil16(nc)/16][Kp/4][64] of A(c, t*nc + c');
  DeviceBuffer<float> A32, bt32;      // bt: band tiles in MFMA lane order [T][n][w][tiles][KS][64]
  // the angular matrix as row lists of its non-zeros (sparse angular form): row c' holds
  // entries [sp_rowptr[c'], sp_rowptr[c' + 1]) of (Z row t nc + c, A_t[c'][c], term t)
  DeviceBuffer<int> sp_rowptr, sp_k, sp_term, sp_pot;  // sp_pot: the entry's potential (time scaling) index
  DeviceBuffer<double> sp_coef64;
  DeviceBuffer<float> sp_coef32;
  mutable DeviceBuffer<double> W64;   // radial products Z [T][nc][n][n] complex
  mutable DeviceBuffer<float> W32;
  mutable DeviceBuffer<double> X64;   // channel-minor copy of the input
  mutable DeviceBuffer<float> X32;
  // Optional overlap term: radial term index nterms holds S1 (x) S2 (identity angular part),
  // added with its own coefficient beta outside the angular contraction.
  bool has_overlap = false;
  // Separable terms R_t = a (x) b run as two 1-D band passes (fham_separable_kernel): terms
  // [0, nband) use band tiles, terms [nband, nterms) and the overlap term the factors
  // sa/sb (band rows [n][2 bw + 1] per term, the overlap's last).
  bool sep_enabled = false;
  int nband = 0;
  DeviceBuffer<double> sa64, sb64;
  DeviceBuffer<float> sa32, sb32;
  // Position tiles of the band terms (with the separable split only): [ceil(nband / 16)][n n]
  // [ceil(w^2 / 4)][64], lane l of step ks holding R_t[i1][jj][d][i2] at 4 ks + l / 16 =
  // jj w + d, t = 16 tile + l % 16
  DeviceBuffer<double> rp64;
  DeviceBuffer<float> rp32;
  // FP32 overlap in Cholesky form (set by the mixed solver from its representation): the
  // upper band factors U1, U2 of the two radial overlaps ([j][csd + i - j]); with them an
  // FP32 multiply with a beta forms (S1 (x) S2) x as U1^T U1 x U2^T U2 in the separable kernel
  // (a cache of the representation's factors, attached by a solver: mutable)
  mutable int csd = -1;
  mutable DeviceBuffer<float> cu1_32, cu2_32;
  void set_cholesky32(const std::vector<double>& u1, const std::vector<double>& u2, int sd) const;
  int rad_variant64 = 0, rad_variant32 = 0;  // radial kernel shape per precision
  int ang_variant64 = 0, ang_variant32 = 0;  // angular stage wave count per precision
  mutable unsigned long long gen64 = 0;      // FP64 multiplies issued
  mutable int w64_terms = 0;                 // radial terms the FP64 products W64 are known to hold
  mutable DeviceBuffer<int> reuse_flag;      // set by the input comparison, re-armed by the angular stage
  // dst = addend + alpha * sum_t scalings[term_pot[t]] (A_t (x) R_t) src + beta (S1 (x) S2) src
  // (beta must be 0 without an overlap term; the overlap radial product is only formed when
  // a beta is non-zero)
  // sc_dev (optional): the same scalings in device memory, [npot] doubles, read by the sparse
  // angular form (so that a captured graph sees each replay's values); the GEMM forms take
  // the host values as kernel arguments
  void apply64(const double* scalings, const double2* src, double2* dst, const double2* addend, double2 alpha,
               hipStream_t s, double2 beta = make_double2(0.0, 0.0), const double* sc_dev = nullptr) const;
  // zs_ext (FP32, may be NULL): the overlap product (S (x) S) src already formed elsewhere (by
  // the Kronecker solve that produced src); it then takes the place of the operator's own
  void apply32(const double* scalings, const float2* src, float2* dst, const float2* addend, double2 alpha,
               hipStream_t s, double2 beta = make_double2(0.0, 0.0), const double* sc_dev = nullptr,
               const float2* zs_ext = nullptr) const;
  bool sparse_angular() const;  // both precisions use a sparse angular form (device scalings honoured)
  // FP64 multiply with up to two outputs from the same products (o2 may be NULL and takes no
  // addend).  reuse_ref (may be NULL): the input of an earlier FP64 multiply whose radial
  // products W64 still holds (covering the terms needed now); the radial stage is skipped when
  // src equals it element for element, checked on the device.
  // npart (may be NULL; FP64 sparse angular form only): per angular workgroup L, npart[2 L]
  // = |o1.dst|^2 and npart[2 L + 1] = |o2->dst|^2 over the block's elements, norm_blocks64()
  // workgroups (the outputs' squared norms without a separate pass)
  void apply64_ex(const double* scalings, const double2* src, const FHamOut& o1, const FHamOut* o2,
                  const double2* reuse_ref, hipStream_t s, const double* sc_dev = nullptr,
                  double* npart = nullptr) const;
  long long norm_blocks64() const;  // 0 when the FP64 angular form is not sparse
};


}  // namespace pp

struct pp_fham {
  pp::FHam f;
};

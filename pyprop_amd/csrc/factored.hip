// Factored Hamiltonian multiply on MFMA (no reference counterpart for the evaluation order;
// the operator is the one the tensor potentials define, generator.py:302-762).
//
//   H X = sum_t f_t (A_t (x) R_t) X,   X: [nc][n][n] complex
//
// Stage 1, angular contraction (the multipole/dipole angular coupling as one GEMM):
//   W[t][c'][q] = f_t sum_c A_t[c'][c] X[c][q]       q over the 2 n^2 reals of a channel
//   M = T nc rows, K = nc, N = 2 n^2; v_mfma_{f32,f64}_16x16x4 with operands read straight
//   from global memory (A is tiny and cache-resident, X rows are read 16 floats at a time).
// Stage 2, radial products, all terms accumulated in registers:
//   Y[c][i1][i2] = sum_t sum_jj sum_d R_t[i1][jj][d][i2] W[t][c][i1-bw+jj][i2-bw+d]
//   One wave per (i1, 16 consecutive i2, 8 channels): per (t, jj) it multiplies the
//   16 x (16+2bw) window of W (rows = 8 channels x re/im) by the dense band tile
//   B[k][m] = R_t[i1][jj][k-m][I+m] (zero outside the band) on MFMA, K = 16+2bw.
//   Epilogue: out = addend + alpha * Y (complex alpha).  src may alias dst or addend: the
//   second stage reads only W and the addend.
#include <hip/hip_runtime.h>

#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"
#include "pp_factored.h"

namespace pp {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <class T>
struct MfmaT;
// Operand layout of both 16x16x4 forms: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16].
// Result: FP32 lane l register r holds D[4 (l / 16) + r][l % 16]; FP64 lane l register r
// holds D[4 r + l / 16][l % 16] (kRowMajorOut below).
template <>
struct MfmaT<float> {
  typedef f32x4 V4;
  typedef float2 C2;
  static constexpr bool kRowMajorOut = true;
  static __device__ inline V4 mma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};
template <>
struct MfmaT<double> {
  typedef f64x4 V4;
  typedef double2 C2;
  static constexpr bool kRowMajorOut = false;
  static __device__ inline V4 mma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};

template <class T>
struct FScales {
  T s[kMaxFTerms];
};

// W[m][q] = scale[m / nc] * sum_k A[m][k] X[k][q]; A is [M][Kp] zero padded, 64 x 64 tile per
// workgroup, four waves of 32 x 32 (2 x 2 MFMA tiles).
template <class T>
__global__ __launch_bounds__(256) void fham_angular_kernel(int M, int nc, int Kp, int Nq, const T* __restrict__ A,
                                                           const T* __restrict__ X, FScales<T> sc, T* __restrict__ W) {
  typedef typename MfmaT<T>::V4 V4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32, n0 = blockIdx.x * 64 + (wave & 1) * 32;
  V4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = V4{0, 0, 0, 0};
  const int lr = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < Kp; k0 += 4) {
    const int k = k0 + lk;
    T a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int row = m0 + i * 16 + lr;
      a[i] = row < M ? A[(long long)row * Kp + k] : T(0);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + j * 16 + lr;
      b[j] = (k < nc && col < Nq) ? X[(long long)k * Nq + col] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = MfmaT<T>::mma(a[i], b[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + i * 16 + (MfmaT<T>::kRowMajorOut ? lk * 4 + r : r * 4 + lk);
        const int col = n0 + j * 16 + lr;
        if (row < M && col < Nq) W[(long long)row * Nq + col] = sc.s[row / nc] * acc[i][j][r];
      }
}

// One wave per (i2 tile of 16, channel group of 8, i1).  W: [T][nc][n][n] complex (as reals),
// dia: [T][n][w][w][n].
template <class T>
__global__ __launch_bounds__(64) void fham_radial_kernel(int nc, int n, int bw, int nterms, int ngroups,
                                                         const T* __restrict__ W, const T* __restrict__ dia,
                                                         const typename MfmaT<T>::C2* __restrict__ addend,
                                                         typename MfmaT<T>::C2 alpha,
                                                         typename MfmaT<T>::C2* __restrict__ out) {
  typedef typename MfmaT<T>::V4 V4;
  typedef typename MfmaT<T>::C2 C2;
  const int lane = threadIdx.x;
  const int tile = blockIdx.x / ngroups, grp = blockIdx.x % ngroups;
  const int i1 = blockIdx.y;
  const int I = tile * 16;
  const int w = 2 * bw + 1;
  const int KW = 16 + 2 * bw;
  const int KS = (KW + 3) >> 2;
  const int lr = lane & 15, lk = lane >> 4;
  // A operand row lr of the wave's 16 rows (8 channels x re/im).  FP32: row = 2 c + part;
  // FP64: row = 8 part + c, so that each lane's four results hold both parts of two
  // channels in either result layout.
  const int ca = grp * 8 + (MfmaT<T>::kRowMajorOut ? (lr >> 1) : (lr & 7));
  const int part = MfmaT<T>::kRowMajorOut ? (lr & 1) : (lr >> 3);
  const bool row_ok = ca < nc;
  const long long nn = (long long)n * n;
  V4 acc = V4{0, 0, 0, 0};
  for (int t = 0; t < nterms; ++t) {
    const T* Wt = W + ((long long)t * nc + (row_ok ? ca : 0)) * nn * 2;
    const T* Dt = dia + (long long)t * n * w * w * n + (long long)i1 * w * w * n;
    for (int jj = 0; jj < w; ++jj) {
      const int j1 = i1 - bw + jj;
      if (j1 < 0 || j1 >= n) continue;  // uniform over the wave
      const T* Wrow = Wt + (long long)j1 * n * 2;
      const T* Dj = Dt + (long long)jj * w * n;
      for (int ks = 0; ks < KS; ++ks) {
        const int k = ks * 4 + lk;
        const int j2 = I - bw + k;
        const T a = (row_ok && k < KW && j2 >= 0 && j2 < n) ? Wrow[2 * j2 + part] : T(0);
        const int d = k - lr;  // B[k][m] with m = lr
        const int i2 = I + lr;
        const T b = (d >= 0 && d < w && i2 < n) ? Dj[(long long)d * n + i2] : T(0);
        acc = MfmaT<T>::mma(a, b, acc);
      }
    }
  }
  // FP32: registers r hold rows 4 lk + r = channels 2 lk, 2 lk + 1 as (re, im, re, im);
  // FP64: registers r hold rows 4 r + lk = channels lk, lk + 4 as (re, re, im, im).
  const int i2 = I + lr;
  if (i2 >= n) return;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = grp * 8 + (MfmaT<T>::kRowMajorOut ? lk * 2 + q : lk + 4 * q);
    if (c >= nc) continue;
    const long long idx = ((long long)c * n + i1) * n + i2;
    const T vr = MfmaT<T>::kRowMajorOut ? acc[2 * q] : acc[q];
    const T vi = MfmaT<T>::kRowMajorOut ? acc[2 * q + 1] : acc[q + 2];
    C2 o;
    if (addend) o = addend[idx];
    else {
      o.x = 0;
      o.y = 0;
    }
    o.x += alpha.x * vr - alpha.y * vi;
    o.y += alpha.x * vi + alpha.y * vr;
    out[idx] = o;
  }
}

template <class T>
static void fham_run(const FHam& F, const double* scalings, const typename MfmaT<T>::C2* src,
                     typename MfmaT<T>::C2* dst, const typename MfmaT<T>::C2* addend, double2 alpha,
                     const DeviceBuffer<T>& Ab, const DeviceBuffer<T>& Db, DeviceBuffer<T>& Wb, hipStream_t s) {
  const int nc = F.nc, n = F.n, nterms = F.nterms;
  const long long Nq = 2LL * n * n;
  if (Wb.count < (size_t)nterms * nc * Nq) Wb.resize((size_t)nterms * nc * Nq);
  FScales<T> sc;
  for (int t = 0; t < kMaxFTerms; ++t) sc.s[t] = t < nterms ? (T)scalings[F.term_pot[t]] : T(0);
  const int M = nterms * nc;
  hipLaunchKernelGGL(fham_angular_kernel<T>, dim3((int)ceil_div(Nq, 64), (int)ceil_div(M, 64)), dim3(256), 0, s, M,
                     nc, F.Kp, (int)Nq, Ab.ptr, reinterpret_cast<const T*>(src), sc, Wb.ptr);
  const int ngroups = (nc + 7) / 8;
  const int tiles = (n + 15) / 16;
  typename MfmaT<T>::C2 a;
  a.x = (T)alpha.x;
  a.y = (T)alpha.y;
  hipLaunchKernelGGL(fham_radial_kernel<T>, dim3(tiles * ngroups, n), dim3(64), 0, s, nc, n, F.bw, nterms, ngroups,
                     Wb.ptr, Db.ptr, addend, a, dst);
  PP_HIP(hipGetLastError());
}

void fham_apply64(const FHam& f, const double* scalings, const double2* src, double2* dst, const double2* addend,
                  double2 alpha, hipStream_t s) {
  f.apply64(scalings, src, dst, addend, alpha, s);
}

void FHam::apply64(const double* scalings, const double2* src, double2* dst, const double2* addend, double2 alpha,
                   hipStream_t s) const {
  fham_run<double>(*this, scalings, src, dst, addend, alpha, A64, dia64, W64, s);
}

void FHam::apply32(const double* scalings, const float2* src, float2* dst, const float2* addend, double2 alpha,
                   hipStream_t s) const {
  fham_run<float>(*this, scalings, src, dst, addend, alpha, A32, dia32, W32, s);
}

}  // namespace pp

using namespace pp;

extern "C" int pp_fham_create(int nchan, int n, int nterms, int bw, const double* A, const int* term_potential,
                              const double* dia, int npotentials, pp_fham** out) {
  return guarded([&] {
    PP_REQUIRE(out && A && term_potential && dia, "null argument");
    PP_REQUIRE(nchan > 0 && n > 0 && bw >= 0 && 2 * bw + 1 <= n, "invalid factored-operator shape");
    PP_REQUIRE(nterms > 0 && nterms <= kMaxFTerms, "factored operator supports 1..32 terms");
    for (int t = 0; t < nterms; ++t)
      PP_REQUIRE(term_potential[t] >= 0 && term_potential[t] < npotentials, "term potential index out of range");
    auto* p = new pp_fham();
    FHam& F = p->f;
    F.nc = nchan;
    F.n = n;
    F.nterms = nterms;
    F.bw = bw;
    F.npot = npotentials;
    F.Kp = (nchan + 3) / 4 * 4;
    F.term_pot.assign(term_potential, term_potential + nterms);
    const size_t M = (size_t)nterms * nchan;
    std::vector<double> a(M * F.Kp, 0.0);
    for (size_t r = 0; r < M; ++r)
      for (int c = 0; c < nchan; ++c) a[r * F.Kp + c] = A[r * nchan + c];
    const int w = 2 * bw + 1;
    const size_t nd = (size_t)nterms * n * w * w * n;
    try {
      hipStream_t s = nullptr;
      F.A64.upload(a.data(), a.size(), s);
      std::vector<float> af(a.begin(), a.end());
      F.A32.upload(af.data(), af.size(), s);
      F.dia64.upload(dia, nd, s);
      std::vector<float> df(dia, dia + nd);
      F.dia32.upload(df.data(), nd, s);
      PP_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

extern "C" int pp_fham_destroy(pp_fham* f) {
  return guarded([&] { delete f; });
}

extern "C" int pp_fham_multiply(const pp_fham* f, const double* scalings, const pp_cplx* src, pp_cplx* dst,
                                const pp_cplx* addend, pp_cplx alpha, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(f && scalings && src && dst, "null argument");
    f->f.apply64(scalings, reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst),
                 reinterpret_cast<const double2*>(addend), make_double2(alpha.re, alpha.im), (hipStream_t)stream);
  });
}

extern "C" int pp_fham_multiply_f32(const pp_fham* f, const double* scalings, const pp_cplxf* src, pp_cplxf* dst,
                                    const pp_cplxf* addend, pp_cplx alpha, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(f && scalings && src && dst, "null argument");
    f->f.apply32(scalings, reinterpret_cast<const float2*>(src), reinterpret_cast<float2*>(dst),
                 reinterpret_cast<const float2*>(addend), make_double2(alpha.re, alpha.im), (hipStream_t)stream);
  });
}

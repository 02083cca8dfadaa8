// Factored Hamiltonian multiply on MFMA (no reference counterpart for the evaluation order;
// the operator is the one the tensor potentials define, generator.py:302-762).
//
//   H X = sum_t f_t (A_t (x) R_t) X,   X: [nc][n][n] complex
//
// Stage 1, radial products (one per term, all channels; the input X is L2-resident):
//   Z[t][c][i1][i2] = sum_jj sum_d R_t[i1][jj][d][i2] X[c][i1-bw+jj][i2-bw+d]
//   One wave per (t, i1, 16 consecutive i2, 8 channels): per jj it multiplies the
//   16 x (16+2bw) window of X (rows = 8 channels x re/im) by the dense band tile
//   B[k][m] = R_t[i1][jj][k-m][I+m] (zero outside the band) on MFMA, K = 16+2bw.
//   Workgroups are renumbered so that the waves sharing (t, i1) -- and so one radial block --
//   run on one XCD and share its L2.
// Stage 2, angular contraction of all terms at once (the multipole and dipole angular
// coupling as one GEMM, K = T nc):
//   Y[c'][p] = addend + alpha sum_{t,c} f_t A_t[c'][c] Z[t][c][p]
//   real A, complex Z: two accumulators (re, im) so each lane holds whole complex outputs.
// Both stages use v_mfma_{f32,f64}_16x16x4.
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

// Channel-minor copy of the input: Xt[j1][j2][c][part] (channels padded to ncp with zeros),
// so that the 16 operand rows of a wave (8 channels x re/im) at one j2 are 16 consecutive
// values.
template <class T>
__global__ void fham_transpose_kernel(int nc, int ncp, int n, const T* __restrict__ X, T* __restrict__ Xt,
                                      const int* __restrict__ skip) {
  if (skip && *skip) return;
  const long long total = (long long)n * n * ncp * 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int part = (int)(i & 1);
    const int c = (int)((i >> 1) % ncp);
    const long long pos = (i >> 1) / ncp;  // j1 * n + j2
    Xt[i] = c < nc ? X[((long long)c * n * n + pos) * 2 + part] : T(0);
  }
}

// Stage 1.  Z: [T][nc][n][n] complex (as reals).  BT: the band tiles in MFMA lane order,
// [T][n][w][tiles][KS][64] (element B[k][m] = R_t[i1][jj][k-m][I+m] at k = 4 ks + lane/16,
// m = lane % 16), so every operand load is one coalesced wave-wide read.
template <class T, int BW>
__global__ __launch_bounds__(64) void fham_radial_kernel(int nc, int ncp, int n, int nterms, int ngroups, int tiles,
                                                         long long nblocks, const T* __restrict__ Xt,
                                                         const T* __restrict__ BT, T* __restrict__ Z,
                                                         const int* __restrict__ skip) {
  if (skip && *skip) return;
  typedef typename MfmaT<T>::V4 V4;
  constexpr int w = 2 * BW + 1;
  constexpr int KW = 16 + 2 * BW;
  constexpr int KS = (KW + 3) / 4;
  // XCD-aware renumbering: hardware block h runs on XCD h % 8; logical blocks
  // [x * per, (x+1) * per) go to XCD x, so consecutive logical blocks share an L2.
  const long long per = (nblocks + 7) / 8;
  const long long hw = blockIdx.x;
  const long long L = (hw % 8) * per + hw / 8;
  if (L >= nblocks) return;
  const int grp = (int)(L % ngroups);
  const int tile = (int)((L / ngroups) % tiles);
  const int i1 = (int)((L / ((long long)ngroups * tiles)) % n);
  const int t = (int)(L / ((long long)ngroups * tiles * n));
  const int lane = threadIdx.x;
  const int I = tile * 16;
  const int lr = lane & 15, lk = lane >> 4;
  // A operand row lr of the wave's 16 rows (8 channels x re/im).  FP32: row = 2 c + part;
  // FP64: row = 8 part + c, so that each lane's four results hold both parts of two
  // channels in either result layout.
  const int cl = MfmaT<T>::kRowMajorOut ? (lr >> 1) : (lr & 7);
  const int part = MfmaT<T>::kRowMajorOut ? (lr & 1) : (lr >> 3);
  const int xoff = (grp * 8 + cl) * 2 + part;
  const int i2 = I + lr;
  const T* Bt = BT + (((long long)t * n + i1) * w * tiles + tile) * KS * 64 + lane;
  const int jlo = max(0, BW - i1), jhi = min(2 * BW, n - 1 - i1 + BW);
  // jj in batches of JB: all 2 KS JB operand loads of a batch are issued before its MFMAs
  constexpr int JB = sizeof(T) == 8 ? 3 : 5;
  V4 acc0 = V4{0, 0, 0, 0}, acc1 = V4{0, 0, 0, 0};
  for (int jb = jlo; jb <= jhi; jb += JB) {
    T av[JB][KS], bv[JB][KS];
#pragma unroll
    for (int u = 0; u < JB; ++u) {
      const int jj = jb + u;
      const bool jok = jj <= jhi;
      const int j1 = i1 - BW + jj;
      const T* Xrow = Xt + (long long)j1 * n * ncp * 2 + xoff;
      const T* Bj = Bt + (long long)jj * tiles * KS * 64;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int k = ks * 4 + lk;
        const int j2 = I - BW + k;
        av[u][ks] = (jok && k < KW && j2 >= 0 && j2 < n) ? Xrow[(long long)j2 * ncp * 2] : T(0);
        bv[u][ks] = jok ? Bj[ks * 64] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < JB; ++u)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        if ((u * KS + ks) & 1) acc1 = MfmaT<T>::mma(av[u][ks], bv[u][ks], acc1);
        else acc0 = MfmaT<T>::mma(av[u][ks], bv[u][ks], acc0);
      }
  }
  const V4 acc = acc0 + acc1;
  if (i2 >= n) return;
  // FP32: registers r hold rows 4 lk + r = channels 2 lk, 2 lk + 1 as (re, im, re, im);
  // FP64: registers r hold rows 4 r + lk = channels lk, lk + 4 as (re, re, im, im).
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = grp * 8 + (MfmaT<T>::kRowMajorOut ? lk * 2 + q : lk + 4 * q);
    if (c >= nc) continue;
    T* z = Z + ((((long long)t * nc + c) * n + i1) * n + i2) * 2;
    z[0] = MfmaT<T>::kRowMajorOut ? acc[2 * q] : acc[q];
    z[1] = MfmaT<T>::kRowMajorOut ? acc[2 * q + 1] : acc[q + 2];
  }
}

// Stage 2.  Workgroup of kAngWaves waves per (16 rows c', 16 complex positions p); the waves
// split K = T nc and reduce through LDS.  A: [Mp][K] (row c', column t*nc + c).
constexpr int kAngWaves = 16;
constexpr int kAngBatch = 12;
template <class T>
__global__ __launch_bounds__(64 * kAngWaves) void fham_angular_kernel(int nc, int K, int Kreal, long long P,
                                                                      const T* __restrict__ A,
                                                                      const typename MfmaT<T>::C2* __restrict__ Z,
                                                                      FScales<T> sc,
                                                                      const typename MfmaT<T>::C2* addend,
                                                                      typename MfmaT<T>::C2 alpha,
                                                                      typename MfmaT<T>::C2* out) {
  typedef typename MfmaT<T>::V4 V4;
  typedef typename MfmaT<T>::C2 C2;
  __shared__ V4 red[kAngWaves - 1][2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int m0 = blockIdx.y * 16;
  const long long p0 = (long long)blockIdx.x * 16;
  const int nks = K >> 2;
  const int per = (nks + kAngWaves - 1) / kAngWaves;
  const int ks0 = wave * per, ks1 = min(nks, ks0 + per);
  const long long pc = p0 + lr;
  const bool col_ok = pc < P;
  V4 are = V4{0, 0, 0, 0}, aim = V4{0, 0, 0, 0};
  for (int kb = ks0; kb < ks1; kb += kAngBatch) {
    T a[kAngBatch];
    C2 z[kAngBatch];
#pragma unroll
    for (int q = 0; q < kAngBatch; ++q) {
      const int ks = kb + q;
      const int k = ks * 4 + lk;
      const bool ok = ks < ks1 && k < Kreal;
      a[q] = ok ? A[(long long)(m0 + lr) * K + k] * sc.s[k / nc] : T(0);
      if (ok && col_ok) z[q] = Z[(long long)k * P + pc];
      else {
        z[q].x = 0;
        z[q].y = 0;
      }
    }
#pragma unroll
    for (int q = 0; q < kAngBatch; ++q) {
      are = MfmaT<T>::mma(a[q], z[q].x, are);
      aim = MfmaT<T>::mma(a[q], z[q].y, aim);
    }
  }
  if (wave > 0) {
    red[wave - 1][0][lane] = are;
    red[wave - 1][1][lane] = aim;
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int q = 0; q < kAngWaves - 1; ++q) {
    are += red[q][0][lane];
    aim += red[q][1][lane];
  }
  if (!col_ok) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int c = m0 + (MfmaT<T>::kRowMajorOut ? lk * 4 + r : r * 4 + lk);
    if (c >= nc) continue;
    const long long idx = (long long)c * P + pc;
    C2 o;
    if (addend) o = addend[idx];
    else {
      o.x = 0;
      o.y = 0;
    }
    const T vr = are[r], vi = aim[r];
    o.x += alpha.x * vr - alpha.y * vi;
    o.y += alpha.x * vi + alpha.y * vr;
    out[idx] = o;
  }
}

template <class T>
static void fham_run(const FHam& F, const double* scalings, const typename MfmaT<T>::C2* src,
                     typename MfmaT<T>::C2* dst, const typename MfmaT<T>::C2* addend, double2 alpha,
                     const DeviceBuffer<T>& Ab, const DeviceBuffer<T>& Db, DeviceBuffer<T>& Zb, DeviceBuffer<T>& Xb,
                     const int* skip_radial, hipStream_t s) {
  const int nc = F.nc, n = F.n, nterms = F.nterms;
  const long long P = (long long)n * n;
  if (Zb.count < (size_t)nterms * nc * P * 2) Zb.resize((size_t)nterms * nc * P * 2);
  const int ngroups = (nc + 7) / 8;
  const int ncp = ngroups * 8;
  if (Xb.count < (size_t)P * ncp * 2) Xb.resize((size_t)P * ncp * 2);
  hipLaunchKernelGGL(fham_transpose_kernel<T>, dim3((unsigned)std::min<long long>(4096, ceil_div(P * ncp * 2, 256))),
                     dim3(256), 0, s, nc, ncp, n, reinterpret_cast<const T*>(src), Xb.ptr, skip_radial);
  const int tiles = (n + 15) / 16;
  const long long nblocks = (long long)nterms * n * tiles * ngroups;
  const long long hw = (nblocks + 7) / 8 * 8;
  switch (F.bw) {
#define PP_RAD(B)                                                                                                    \
  case B:                                                                                                            \
    hipLaunchKernelGGL((fham_radial_kernel<T, B>), dim3((unsigned)hw), dim3(64), 0, s, nc, ncp, n, nterms, ngroups,   \
                       tiles, nblocks, Xb.ptr, Db.ptr, Zb.ptr, skip_radial);                                          \
    break;
    PP_RAD(1) PP_RAD(2) PP_RAD(3) PP_RAD(4) PP_RAD(5) PP_RAD(6) PP_RAD(7) PP_RAD(8)
#undef PP_RAD
    default:
      throw Error(PP_ERR_INVALID, "factored multiply supports band half-widths 1..8");
  }
  FScales<T> sc;
  for (int t = 0; t < kMaxFTerms; ++t) sc.s[t] = t < nterms ? (T)scalings[F.term_pot[t]] : T(0);
  typename MfmaT<T>::C2 a;
  a.x = (T)alpha.x;
  a.y = (T)alpha.y;
  hipLaunchKernelGGL(fham_angular_kernel<T>, dim3((unsigned)ceil_div(P, 16), (unsigned)ceil_div(nc, 16)),
                     dim3(64 * kAngWaves),
                     0, s, nc, F.Kp, nterms * nc, P, Ab.ptr, reinterpret_cast<const typename MfmaT<T>::C2*>(Zb.ptr), sc, addend, a,
                     dst);
  PP_HIP(hipGetLastError());
}

void fham_apply64(const FHam& f, const double* scalings, const double2* src, double2* dst, const double2* addend,
                  double2 alpha, hipStream_t s) {
  f.apply64(scalings, src, dst, addend, alpha, s);
}

void FHam::apply64(const double* scalings, const double2* src, double2* dst, const double2* addend, double2 alpha,
                   hipStream_t s, const int* skip_radial) const {
  ++gen64;
  fham_run<double>(*this, scalings, src, dst, addend, alpha, A64, bt64, W64, X64, skip_radial, s);
}

// flag = 0 when a and b hold the same values, 1 otherwise (flag preset to 0 by the first
// block's thread 0 is not ordered with the others, so the caller presets it)
__global__ void equal_check_kernel(long long n, const double2* __restrict__ a, const double2* __restrict__ b,
                                   int* __restrict__ differs) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    if (a[i].x != b[i].x || a[i].y != b[i].y) *differs = 1;
}
__global__ void set_flag_kernel(int* __restrict__ f, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *f = v;
}

void fham_flag_differs(long long n, const double2* a, const double2* b, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(64), 0, s, flag, 0);
  hipLaunchKernelGGL(equal_check_kernel, dim3((unsigned)std::min<long long>(2048, ceil_div(n, 256))), dim3(256), 0, s, n,
                     a, b, flag);
  PP_HIP(hipGetLastError());
}

void FHam::apply32(const double* scalings, const float2* src, float2* dst, const float2* addend, double2 alpha,
                   hipStream_t s) const {
  fham_run<float>(*this, scalings, src, dst, addend, alpha, A32, bt32, W32, X32, nullptr, s);
}

}  // namespace pp

using namespace pp;

extern "C" int pp_fham_create(int nchan, int n, int nterms, int bw, const double* A, const int* term_potential,
                              const double* dia, int npotentials, pp_fham** out) {
  return guarded([&] {
    PP_REQUIRE(out && A && term_potential && dia, "null argument");
    PP_REQUIRE(nchan > 0 && nchan <= 64 && n > 0 && bw >= 1 && bw <= 8 && 2 * bw + 1 <= n,
               "factored operator supports 1..64 channels and band half-widths 1..8");
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
    F.Kp = (nterms * nchan + 3) / 4 * 4;  // angular GEMM K, zero-padded to the MFMA K step
    F.term_pot.assign(term_potential, term_potential + nterms);
    const int Mp = (nchan + 15) / 16 * 16;
    // A[c'][t*nc + c] = A_t[c'][c], rows padded to a multiple of 16 with zeros
    std::vector<double> a((size_t)Mp * F.Kp, 0.0);
    for (int t = 0; t < nterms; ++t)
      for (int r = 0; r < nchan; ++r)
        for (int c = 0; c < nchan; ++c) a[(size_t)r * F.Kp + (size_t)t * nchan + c] = A[((size_t)t * nchan + r) * nchan + c];
    const int w = 2 * bw + 1;
    const int tiles = (n + 15) / 16;
    const int KS = (16 + 2 * bw + 3) / 4;
    // band tiles in MFMA lane order: [t][i1][jj][tile][ks][lane]
    const size_t nbt = (size_t)nterms * n * w * tiles * KS * 64;
    std::vector<double> bt(nbt, 0.0);
    for (int t = 0; t < nterms; ++t)
      for (int i1 = 0; i1 < n; ++i1)
        for (int jj = 0; jj < w; ++jj)
          for (int tile = 0; tile < tiles; ++tile)
            for (int ks = 0; ks < KS; ++ks)
              for (int lane = 0; lane < 64; ++lane) {
                const int k = ks * 4 + lane / 16, m = lane % 16;
                const int d = k - m, i2 = tile * 16 + m;
                if (d < 0 || d >= w || i2 >= n) continue;
                bt[((((((size_t)t * n + i1) * w + jj) * tiles + tile) * KS + ks) * 64) + lane] =
                    dia[((((size_t)t * n + i1) * w + jj) * w + d) * n + i2];
              }
    try {
      hipStream_t s = nullptr;
      F.A64.upload(a.data(), a.size(), s);
      std::vector<float> af(a.begin(), a.end());
      F.A32.upload(af.data(), af.size(), s);
      F.bt64.upload(bt.data(), nbt, s);
      std::vector<float> bf(bt.begin(), bt.end());
      F.bt32.upload(bf.data(), nbt, s);
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

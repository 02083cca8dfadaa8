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

#include <algorithm>
#include <cstdlib>
#include <type_traits>
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
// values.  With ref (the input of the multiply whose radial products are held), any element
// where X differs from ref sets *differs = 1; the radial stage then recomputes, and skips
// when *differs stayed 0.
template <class T>
__global__ void fham_transpose_kernel(int nc, int ncp, int n, const T* __restrict__ X, T* __restrict__ Xt,
                                      const T* __restrict__ ref, int* __restrict__ differs) {
  const long long total = (long long)n * n * ncp * 2;
  bool diff = false;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int part = (int)(i & 1);
    const int c = (int)((i >> 1) % ncp);
    const long long pos = (i >> 1) / ncp;  // j1 * n + j2
    const long long src = ((long long)c * n * n + pos) * 2 + part;
    const T v = c < nc ? X[src] : T(0);
    Xt[i] = v;
    if (ref && c < nc && v != ref[src]) diff = true;
  }
  if (diff) *differs = 1;
}

// Stage 1.  Z: [T][nc][n][n] complex (as reals).  BT: the band tiles in MFMA lane order,
// [T][n][w][tiles][KS][64] (element B[k][m] = R_t[i1][jj][k-m][I+m] at k = 4 ks + lane/16,
// m = lane % 16), so every operand load is one coalesced wave-wide read.
template <class T, int BW>
__global__ __launch_bounds__(64) void fham_radial_kernel(int nc, int ncp, int n, int nterms, int ngroups, int tiles,
                                                         long long nblocks, const T* __restrict__ Xt,
                                                         const T* __restrict__ BT, T* __restrict__ Z,
                                                         const int* __restrict__ differs) {
  if (differs && *differs == 0) return;  // input equals the one whose products Z holds
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

// Stage 1 with register reuse: one wave computes TG terms x GG channel groups of one
// (i1, 16-column tile).  Per jj it loads the X windows of its GG groups once and reuses them
// for the band tiles of its TG terms, and each band tile for its GG groups: (GG + TG) KS
// operand loads feed TG GG KS MFMAs (the one-term, one-group form above loads two operands
// per MFMA).  PIPE: the operands of jj + 1 are loaded into a second register set while the
// MFMAs of jj run.  Padding terms (t >= nterms) reuse the last term's tile and are not
// stored; padding groups read zeros and are not stored.
template <class T, int BW, int TG, int GG, bool PIPE>
__global__ __launch_bounds__(64) void fham_radial_reuse_kernel(int nc, int ncp, int n, int nterms, int ngroups,
                                                               int tiles, long long nblocks, const T* __restrict__ Xt,
                                                               const T* __restrict__ BT, T* __restrict__ Z,
                                                               const int* __restrict__ differs) {
  if (differs && *differs == 0) return;  // input equals the one whose products Z holds
  typedef typename MfmaT<T>::V4 V4;
  constexpr int w = 2 * BW + 1;
  constexpr int KW = 16 + 2 * BW;
  constexpr int KS = (KW + 3) / 4;
  const long long per = (nblocks + 7) / 8;
  const long long hw = blockIdx.x;
  const long long L = (hw % 8) * per + hw / 8;
  if (L >= nblocks) return;
  const int ngc = (ngroups + GG - 1) / GG, ntc = (nterms + TG - 1) / TG;
  const int gc = (int)(L % ngc);
  const int tc = (int)((L / ngc) % ntc);
  const int tile = (int)((L / ((long long)ngc * ntc)) % tiles);
  const int i1 = (int)(L / ((long long)ngc * ntc * tiles));
  const int lane = threadIdx.x;
  const int I = tile * 16;
  const int lr = lane & 15, lk = lane >> 4;
  const int cl = MfmaT<T>::kRowMajorOut ? (lr >> 1) : (lr & 7);
  const int part = MfmaT<T>::kRowMajorOut ? (lr & 1) : (lr >> 3);
  const int jlo = max(0, BW - i1), jhi = min(2 * BW, n - 1 - i1 + BW);
  V4 acc[TG][GG];
#pragma unroll
  for (int u = 0; u < TG; ++u)
#pragma unroll
    for (int g = 0; g < GG; ++g) acc[u][g] = V4{0, 0, 0, 0};
  const T* Bt[TG];
#pragma unroll
  for (int u = 0; u < TG; ++u) {
    const int t = min(tc * TG + u, nterms - 1);
    Bt[u] = BT + (((long long)t * n + i1) * w * tiles + tile) * KS * 64 + lane;
  }
  // X window of the GG groups at jj: unconditional loads from clamped (in-bounds) addresses,
  // then a select, so no operand load sits behind a branch
  auto load_x = [&](int jj, T (&av)[GG][KS]) {
    const T* Xrow = Xt + (long long)(i1 - BW + jj) * n * ncp * 2 + part;
#pragma unroll
    for (int g = 0; g < GG; ++g) {
      const int grp = gc * GG + g;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int k = ks * 4 + lk;
        const int j2 = I - BW + k;
        const bool ok = grp < ngroups && k < KW && j2 >= 0 && j2 < n;
        const T v = Xrow[ok ? (long long)j2 * ncp * 2 + (grp * 8 + cl) * 2 : 0];
        av[g][ks] = ok ? v : T(0);
      }
    }
  };
  auto load_b = [&](int jj, int u, T (&bv)[KS]) {
    const T* Bj = Bt[u] + (long long)jj * tiles * KS * 64;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) bv[ks] = Bj[ks * 64];
  };
  auto mma_all = [&](const T (&av)[GG][KS], int u, const T (&bv)[KS]) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int g = 0; g < GG; ++g) acc[u][g] = MfmaT<T>::mma(av[g][ks], bv[ks], acc[u][g]);
  };
  if constexpr (PIPE) {
    T av[GG][KS], bv[TG][KS];
    load_x(jlo, av);
#pragma unroll
    for (int u = 0; u < TG; ++u) load_b(jlo, u, bv[u]);
    for (int jj = jlo; jj <= jhi; ++jj) {
      const int jn = min(jj + 1, jhi);  // the last pass reloads jhi (in bounds, unused)
      T an[GG][KS], bn[TG][KS];
      load_x(jn, an);
#pragma unroll
      for (int u = 0; u < TG; ++u) load_b(jn, u, bn[u]);
      // keep the next step's loads issued here, ahead of this step's MFMAs (the scheduler
      // would otherwise sink them towards their first use)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < TG; ++u) mma_all(av, u, bv[u]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int g = 0; g < GG; ++g)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) av[g][ks] = an[g][ks];
#pragma unroll
      for (int u = 0; u < TG; ++u)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) bv[u][ks] = bn[u][ks];
    }
  } else {
    for (int jj = jlo; jj <= jhi; ++jj) {
      T av[GG][KS];
      load_x(jj, av);
#pragma unroll
      for (int u = 0; u < TG; ++u) {
        T bv[KS];
        load_b(jj, u, bv);
        mma_all(av, u, bv);
      }
    }
  }
  const int i2 = I + lr;
  if (i2 >= n) return;
#pragma unroll
  for (int u = 0; u < TG; ++u) {
    const int t = tc * TG + u;
    if (t >= nterms) continue;
#pragma unroll
    for (int g = 0; g < GG; ++g) {
      const int grp = gc * GG + g;
      if (grp >= ngroups) continue;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c = grp * 8 + (MfmaT<T>::kRowMajorOut ? lk * 2 + q : lk + 4 * q);
        if (c >= nc) continue;
        T* z = Z + ((((long long)t * nc + c) * n + i1) * n + i2) * 2;
        z[0] = MfmaT<T>::kRowMajorOut ? acc[u][g][2 * q] : acc[u][g][q];
        z[1] = MfmaT<T>::kRowMajorOut ? acc[u][g][2 * q + 1] : acc[u][g][q + 2];
      }
    }
  }
}

// Stage 1, position form.  At one position (i1, i2) the radial products of all band terms
// are one small GEMM: Z_t[(c, part)][i1][i2] = sum_k X[(c, part)][k] R[k][t] over the
// w x w window k = jj w + d, X[.][k] = X[c][i1 - BW + jj][i2 - BW + d] (zero outside the
// grid), R[k][t] = dia_t[i1][jj][d][i2].  No band padding enters the MFMAs (the band-tile
// form multiplies (16 + 2 BW)-wide tiles of which w diagonals are non-zero).  RP: [term
// tile of 16][position i1 n + i2][KP][64] in MFMA lane order, lane l of step ks holding
// R[4 ks + l / 16][16 tile + l % 16], zero past w^2 and past the last term.  A wave forms one
// position's products for MG channel groups of 8 (16 operand rows each), sharing each R
// operand among them.  The R operands of KC steps are loaded at once (a wave's R stream
// comes from HBM, so it waits once per chunk instead of once per step batch); the X
// operands, read through L2, are loaded KB steps ahead of their MFMAs.  Waves: (term tile,
// position, group chunk), group chunk fastest, four per workgroup.
template <class T, int BW, int MG>
__global__ __launch_bounds__(256) void fham_radial_pos_kernel(int nc, int ncp, int n, int nterms, int ngroups,
                                                              long long nwaves, const T* __restrict__ Xt,
                                                              const T* __restrict__ RP, T* __restrict__ Z,
                                                              const int* __restrict__ differs) {
  if (differs && *differs == 0) return;  // input equals the one whose products Z holds
  typedef typename MfmaT<T>::V4 V4;
  typedef typename MfmaT<T>::C2 C2;
  constexpr int w = 2 * BW + 1, KK = w * w, KP = (KK + 3) / 4;
  constexpr int KB = 4;                  // steps per X operand batch
  constexpr int KC = sizeof(T) == 8 ? 24 : 44;  // steps per R operand chunk (a multiple of KB)
  constexpr int NA = MG >= 4 ? 1 : 2;    // independent accumulator chains per group
  const long long nblocks = (nwaves + 3) / 4;
  const long long per = (nblocks + 7) / 8;
  const long long hw = blockIdx.x;
  const long long L = (hw % 8) * per + hw / 8;  // XCD-aware: consecutive blocks share an L2
  const long long wv = L * 4 + (threadIdx.x >> 6);
  if (L >= nblocks || wv >= nwaves) return;  // whole waves; the kernel has no barrier
  const int nmg = (ngroups + MG - 1) / MG;
  const long long P = (long long)n * n;
  const int mgc = (int)(wv % nmg);
  const long long pos = (wv / nmg) % P;
  const int tn = (int)(wv / ((long long)nmg * P));
  const int i1 = (int)(pos / n), i2 = (int)(pos - (long long)i1 * n);
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int cl = MfmaT<T>::kRowMajorOut ? (lr >> 1) : (lr & 7);
  const int part = MfmaT<T>::kRowMajorOut ? (lr & 1) : (lr >> 3);
  int roff[MG];  // this lane's operand row within a position's ncp * 2 values, per group
#pragma unroll
  for (int g = 0; g < MG; ++g) roff[g] = (min(mgc * MG + g, ngroups - 1) * 8 + cl) * 2 + part;
  const T* Rb = RP + ((long long)tn * P + pos) * KP * 64 + lane;
  V4 acc[NA][MG];
#pragma unroll
  for (int h = 0; h < NA; ++h)
#pragma unroll
    for (int g = 0; g < MG; ++g) acc[h][g] = V4{0, 0, 0, 0};
  // X operands, step after step (k = 4 ks + lk = jj w + d advances by 4 per step): the
  // lane's window position (jj, d) and its offset in Xt are carried from step to step.
  // Outside the grid (or past k = w^2, where jj reaches w) the load reads the position-0
  // element and the select zeroes it.
  const int S2 = ncp * 2, S1 = n * S2;  // Xt strides of j2 and j1 (Xt has n n ncp 2 < 2^31 values)
  const int jlo = max(0, BW - i1), jhi = min(w - 1, n - 1 - i1 + BW);
  const int dlo = max(0, BW - i2), dhi = min(w - 1, n - 1 - i2 + BW);
  int jj = lk / w, d = lk - (lk / w) * w;  // k = lk at step 0 (w = 3 puts lk = 3 in row 1)
  int off = (i1 - BW + jj) * S1 + (i2 - BW + d) * S2;
  auto load_x = [&](T (&a)[MG]) {
    const bool ok = jj >= jlo && jj <= jhi && d >= dlo && d <= dhi;
    const int o = ok ? off : 0;
#pragma unroll
    for (int g = 0; g < MG; ++g) {
      const T v = Xt[o + roff[g]];
      a[g] = ok ? v : T(0);
    }
    d += 4;
    off += 4 * S2;
    const bool wrap = d >= w;
    d = wrap ? d - w : d;
    jj = wrap ? jj + 1 : jj;
    off = wrap ? off + S1 - w * S2 : off;
    if constexpr (w < 4) {  // w = 3: a step of 4 can cross two rows
      const bool wrap2 = d >= w;
      d = wrap2 ? d - w : d;
      jj = wrap2 ? jj + 1 : jj;
      off = wrap2 ? off + S1 - w * S2 : off;
    }
  };
  T a0[KB][MG];
#pragma unroll
  for (int u = 0; u < KB; ++u) load_x(a0[u]);  // steps 0 .. KB - 1
  for (int c0 = 0; c0 < KP; c0 += KC) {
    T bc[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) bc[u] = Rb[(long long)min(c0 + u, KP - 1) * 64];  // past the end: unused
#pragma unroll
    for (int kb = 0; kb < KC; kb += KB) {
      T a1[KB][MG];
#pragma unroll
      for (int u = 0; u < KB; ++u) load_x(a1[u]);  // steps c0 + kb + KB + u (past KP: zeros, unused)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (c0 + kb + u < KP) {
#pragma unroll
          for (int g = 0; g < MG; ++g) acc[u % NA][g] = MfmaT<T>::mma(a0[u][g], bc[kb + u], acc[u % NA][g]);
        }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < KB; ++u)
#pragma unroll
        for (int g = 0; g < MG; ++g) a0[u][g] = a1[u][g];
    }
  }
  const int t = tn * 16 + lr;  // result column: the term
  if (t >= nterms) return;
#pragma unroll
  for (int g = 0; g < MG; ++g) {
    const int grp = mgc * MG + g;
    if (grp >= ngroups) continue;
    V4 r = acc[0][g];
    if constexpr (NA == 2) r = r + acc[1][g];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = grp * 8 + (MfmaT<T>::kRowMajorOut ? lk * 2 + q : lk + 4 * q);
      if (c >= nc) continue;
      C2 v;
      v.x = MfmaT<T>::kRowMajorOut ? r[2 * q] : r[q];
      v.y = MfmaT<T>::kRowMajorOut ? r[2 * q + 1] : r[q + 2];
      reinterpret_cast<C2*>(Z)[(((long long)t * nc + c) * n + i1) * n + i2] = v;
    }
  }
}

// LDS layout of fham_radial_win_kernel: values per window point (FP32: an odd multiple of 16,
// so that the two points a half-wave reads lie in opposite halves of the 32 banks) and bytes.
// SW (FP32, even MG): no padding; instead value v of point p sits at p PS + (v ^ 16 (p & 1)),
// which puts consecutive points' 16-value halves in opposite bank halves just the same, in
// the smaller window that lets three workgroups share a CU.
template <class T, int MG, bool SW = false>
__host__ __device__ constexpr int win_point_stride() {
  return sizeof(T) == 4 && MG % 2 == 0 && !SW ? MG * 16 + 16 : MG * 16;
}
// the largest divisor of rows that is at most cap (rows of the window copied per batch)
__host__ __device__ constexpr int win_rows_per_batch(int rows, int cap) {
  return cap <= 1 ? 1 : rows % cap == 0 ? cap : win_rows_per_batch(rows, cap - 1);
}
// LDS of fham_radial_win_kernel with nw waves: the window, then 16 bytes per thread that the
// copy's slots outside the window write; afterwards the same space holds the products,
// MG x 8 x 16 rows of TW + 1 complex values
__host__ __device__ constexpr size_t win_lds_bytes(int bw, int mg, int tw, int nw, int esz, bool sw = false) {
  const size_t win = (size_t)(2 * bw + 2) * (tw + 2 * bw) * (esz == 4 && mg % 2 == 0 && !sw ? mg * 16 + 16 : mg * 16) *
                         esz +
                     (size_t)64 * nw * 16;
  const size_t out = (size_t)mg * 8 * 16 * (tw + 1) * 2 * esz;
  return win > out ? win : out;
}
template <class T, int BW, int MG, int TW, int NW, bool SW = false>
__host__ __device__ constexpr size_t win_lds_bytes() {
  return win_lds_bytes(BW, MG, TW, NW, (int)sizeof(T), SW);
}

// Stage 1, position form with the window in LDS.  A workgroup takes one r1 point i1, TW
// consecutive r2 points i0 .. i0 + TW - 1 and channel groups gc MG .. gc MG + MG - 1.  It
// copies the X values of the union of their windows from Xt into LDS once; its NW waves
// then form the products of TW / NW positions each, the GEMM of fham_radial_pos_kernel per
// position, with the X operands read from LDS and the next position's R operands loaded
// while the current position's MFMAs run.  LDS point (jr, jc), jr <= w, jc < TW + 2 BW,
// holds the MG x 16 operand values of grid point (i1 - BW + jr, i0 - BW + jc): zero outside
// the grid, and in the extra row jr = w, which the last K step's lanes past w^2 read (their
// R entries are zero).  K step ks, lane (lr, lk): k = 4 ks + lk = w jj + d reads point
// (jj, ip + d), ip the position's offset in the tile.  With (jj0, d0) = divmod(4 ks, w) that
// point is jj0 WC + d0 + lk, plus WC - w where d0 + lk >= w (the row wraps): a lane base
// chosen by ks at compile time plus an immediate, so the reads take no address arithmetic.
template <class T, int BW, int MG, int TW, int NW, bool SW = false>
__global__ __launch_bounds__(64 * NW) void fham_radial_win_kernel(int nc, int n, int nterms, int ngroups, int ntiles,
                                                              long long nblocks, int ncp, const T* __restrict__ Xt,
                                                              const T* __restrict__ RP, T* __restrict__ Z,
                                                              const int* __restrict__ differs) {
  if (differs && *differs == 0) return;  // input equals the one whose products Z holds
  typedef typename MfmaT<T>::V4 V4;
  typedef typename MfmaT<T>::C2 C2;
  constexpr int w = 2 * BW + 1, KK = w * w, KP = (KK + 3) / 4;
  constexpr int WC = TW + 2 * BW;  // window columns
  constexpr int PS = win_point_stride<T, MG, SW>();
  constexpr int NT = 64 * NW;   // threads
  constexpr int PPW = TW / NW;  // positions per wave
  constexpr int KB = 4;         // K steps per batch of LDS reads
  static_assert(TW % NW == 0, "the waves share the tile's positions evenly");
  static_assert(!SW || (sizeof(T) == 4 && MG % 2 == 0 && PPW % 2 == 0 && (TW - 1) % 2 == 1),
                "the swizzled window: FP32, pairs of 16-value halves, an even first point per wave");
  extern __shared__ __align__(16) char win_raw[];
  T* win = reinterpret_cast<T*>(win_raw);
  const int per = (int)((nblocks + 7) / 8);  // nblocks < 2^31 (n <= 64)
  const int hw = blockIdx.x;
  const int L = (hw % 8) * per + hw / 8;  // XCD-aware: a tile's chunks share an L2
  if (L >= nblocks) return;               // the whole workgroup, before its barrier
  const int nchunks = (ngroups + MG - 1) / MG;
  const int gc = L % nchunks;
  int rest = L / nchunks;
  const int tile = rest % ntiles;
  rest /= ntiles;
  const int i1 = rest % n;
  const int tn = rest / n;
  const int i0 = tile * TW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long P = (long long)n * n;
  const T* Rt = RP + (long long)tn * P * KP * 64 + lane;
  const int p0 = i0 + wave * PPW;  // this wave's first position
  // R operands: position s of the wave in rb[s % 2] (past n: the last position's, unused)
  T rb[2][KP];
  {  // issued first: their latency overlaps the window copy
    const T* r = Rt + ((long long)i1 * n + min(p0, n - 1)) * KP * 64;
#pragma unroll
    for (int ks = 0; ks < KP; ++ks) rb[0][ks] = r[ks * 64];
  }
  // window copy in 16-byte pieces: thread slot q of column jcb + CPI cs, row by row, RB rows
  // per batch whose loads are all issued before their LDS stores.  Buffer loads: a piece
  // outside the grid (or past the channel groups) gets an offset beyond the buffer's range,
  // for which the load returns zeros, so no load sits behind a branch.
  {
    typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type Piece;
    constexpr int VPP = 16 / sizeof(T);  // values per piece
    constexpr int PPP = MG * 16 / VPP;   // pieces per point
    constexpr int SLOTS = PPP <= 4 ? 4 : PPP <= 8 ? 8 : PPP <= 16 ? 16 : 32;
    constexpr int CPI = NT / SLOTS;      // columns per pass
    constexpr int NCS = (WC + CPI - 1) / CPI;
    constexpr int RB = win_rows_per_batch(2 * BW + 2, NCS >= 28 ? 1 : 28 / NCS);
    static_assert(PPP <= 32, "at most 32 pieces per point");
    constexpr int kOut = 0x7ffffff0;     // beyond any buffer here (Xt < 2 GiB)
    const int xt_bytes = n * n * ncp * 2 * (int)sizeof(T);
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc((void*)Xt, (short)0, xt_bytes, 0x00020000);
    const int q = (int)threadIdx.x % SLOTS, jcb = (int)threadIdx.x / SLOTS;
    const int spare = (2 * BW + 2) * WC * PS + (int)threadIdx.x * VPP;  // this thread's 16 spare bytes
    const bool qok = q < PPP && gc * MG + q * VPP / 16 < ngroups;
    const int qoff = (gc * MG * 16 + q * VPP) * (int)sizeof(T);
    int coff[NCS];
#pragma unroll
    for (int cs = 0; cs < NCS; ++cs) {
      const int jc = jcb + cs * CPI, j2 = i0 - BW + jc;
      coff[cs] = qok && jc < WC && j2 >= 0 && j2 < n ? j2 * ncp * 2 * (int)sizeof(T) + qoff : -1;
    }
#pragma unroll
    for (int jr0 = 0; jr0 <= w; jr0 += RB) {
      Piece v[RB][NCS];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int jr = jr0 + r, j1 = i1 - BW + jr;
        const bool rok = jr < w && j1 >= 0 && j1 < n;
        const int roff = j1 * n * ncp * 2 * (int)sizeof(T);
#pragma unroll
        for (int cs = 0; cs < NCS; ++cs) {
          const int off = rok && coff[cs] >= 0 ? roff + coff[cs] : kOut;
          const auto x = __builtin_amdgcn_raw_buffer_load_b128(xr, off, 0, 0);
          v[r][cs] = *reinterpret_cast<const Piece*>(&x);
        }
      }
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int jr = jr0 + r;
#pragma unroll
        for (int cs = 0; cs < NCS; ++cs) {  // unconditional stores: no load is sunk into a branch
          const int jc = jcb + cs * CPI;
          const int pt = jr * WC + jc;
          const int dst = jc < WC && q < PPP ? pt * PS + (SW ? (q * VPP) ^ ((pt & 1) << 4) : q * VPP) : spare;
          *reinterpret_cast<Piece*>(win + dst) = v[r][cs];
        }
      }
    }
  }
  __syncthreads();
  const int lr = lane & 15, lk = lane >> 4;
  const int cl = MfmaT<T>::kRowMajorOut ? (lr >> 1) : (lr & 7);
  const int part = MfmaT<T>::kRowMajorOut ? (lr & 1) : (lr >> 3);
  const int b0 = lk * PS + cl * 2 + part;  // point lk of row 0, this lane's operand row
  const int bwr = b0 + (WC - w) * PS;      // the same point after a row wrap
  const int b1 = lk >= 1 ? bwr : b0, b2 = lk >= 2 ? bwr : b0, b3 = lk >= 3 ? bwr : b0;
  // swizzled window: the lane's point is the read's compile-time point offset (even wave
  // start) plus lk, or lk + WC - w (= TW - 1, odd) past a row wrap; eb[h][j] is base j with
  // the 16-value half h ^ (parity of the lane's part of the point) selected
  int eb[2][4];
  if (SW) {
    const int pa = lk & 1, pw = (lk + WC - w) & 1;
    const int par[4] = {pa, lk >= 1 ? pw : pa, lk >= 2 ? pw : pa, lk >= 3 ? pw : pa};
    const int bs[4] = {b0, b1, b2, b3};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      eb[0][j] = bs[j] + (par[j] << 4);
      eb[1][j] = bs[j] + ((par[j] ^ 1) << 4);
    }
  }
  // The wave's positions run as one pipeline: each batch of KB K steps issues the LDS reads
  // of the next batch (the next position's first batch after a position's last) and the next
  // position's R operands with the same K indices, then its MFMAs.  In a wave with a valid
  // position the positions past n are formed too (with the last position's R) and discarded;
  // a wave with none skips the products.
  V4 acc[PPW][MG];
  const bool active = p0 < n;  // wave-uniform
  if (active) {
    const T* wv = win + wave * PPW * PS;  // the wave's first position
    auto read_x = [&](int sp, int ks, T (&a)[MG]) {
      const int d0 = (4 * ks) % w, th = w - d0;  // lanes lk >= th wrap
      const int cpt = (4 * ks) / w * WC + d0 + sp;  // the read's point offset
      const int c0 = cpt * PS;
      if (SW) {
        const int j = th >= 4 ? 0 : th == 3 ? 3 : th == 2 ? 2 : 1;
#pragma unroll
        for (int g = 0; g < MG; ++g) a[g] = wv[eb[(g & 1) ^ (cpt & 1)][j] + c0 + (g & ~1) * 16];
      } else {
        const int base = th >= 4 ? b0 : th == 3 ? b3 : th == 2 ? b2 : b1;
#pragma unroll
        for (int g = 0; g < MG; ++g) a[g] = wv[base + c0 + g * 16];
      }
    };
    const T* rpos[PPW];  // R operands of each position (past n: the last position's)
#pragma unroll
    for (int sp = 0; sp < PPW; ++sp) rpos[sp] = Rt + ((long long)i1 * n + min(p0 + sp, n - 1)) * KP * 64;
#pragma unroll
    for (int sp = 0; sp < PPW; ++sp)
#pragma unroll
      for (int g = 0; g < MG; ++g) acc[sp][g] = V4{0, 0, 0, 0};
    T a0[KB][MG];
#pragma unroll
    for (int u = 0; u < KB; ++u) read_x(0, u, a0[u]);
#pragma unroll
    for (int sp = 0; sp < PPW; ++sp) {
#pragma unroll
      for (int kb = 0; kb < KP; kb += KB) {
        const bool last = kb + KB >= KP;
        const int nsp = last ? sp + 1 : sp, nkb = last ? 0 : kb + KB;  // the next batch
        T a1[KB][MG];
#pragma unroll
        for (int u = 0; u < KB; ++u)
          if (nsp < PPW && nkb + u < KP) read_x(nsp, nkb + u, a1[u]);
        if (sp + 1 < PPW) {
#pragma unroll
          for (int u = 0; u < KB; ++u)
            if (kb + u < KP) rb[(sp + 1) % 2][kb + u] = rpos[sp + 1][(kb + u) * 64];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < KB; ++u)
          if (kb + u < KP) {
#pragma unroll
            for (int g = 0; g < MG; ++g) acc[sp][g] = MfmaT<T>::mma(a0[u][g], rb[sp % 2][kb + u], acc[sp][g]);
          }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < KB; ++u)
#pragma unroll
          for (int g = 0; g < MG; ++g) a0[u][g] = a1[u][g];
      }
    }
  }
  // The products leave through LDS (the window's space once every wave is done with it), as
  // rows (channel, term) of TW consecutive r2 points, so that the global stores are whole
  // rows.  Layout [g][c8][t][OLD]: OLD odd, so that the 16 terms a wave writes at once fall
  // in distinct banks.
  constexpr int OLD = TW + 1;
  __syncthreads();
  C2* outs = reinterpret_cast<C2*>(win_raw);
  if (active) {
#pragma unroll
    for (int sp = 0; sp < PPW; ++sp)
#pragma unroll
      for (int g = 0; g < MG; ++g)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int c8 = MfmaT<T>::kRowMajorOut ? lk * 2 + q : lk + 4 * q;
          C2 v;
          v.x = MfmaT<T>::kRowMajorOut ? acc[sp][g][2 * q] : acc[sp][g][q];
          v.y = MfmaT<T>::kRowMajorOut ? acc[sp][g][2 * q + 1] : acc[sp][g][q + 2];
          outs[((g * 8 + c8) * 16 + lr) * OLD + wave * PPW + sp] = v;
        }
  }
  __syncthreads();
  {
    // thread: r2 point ip and term t fixed, channel rows c8g = c8g0 + RS k (NT / TW rows per
    // pass, a multiple of 16, so t stays the thread's)
    constexpr int RS = NT / TW / 16;
    static_assert(NT % (16 * TW) == 0, "the copy's passes keep each thread's term");
    const int ip = (int)threadIdx.x % TW, t = ((int)threadIdx.x / TW) % 16, c8g0 = (int)threadIdx.x / TW / 16;
    const int tg = tn * 16 + t, i2 = i0 + ip;
    const bool ok = tg < nterms && i2 < n;
    C2* zr = reinterpret_cast<C2*>(Z) + ((long long)tg * nc * n + i1) * n + i2;  // channel 0's row
#pragma unroll
    for (int k = 0; k < MG * 8 / RS; ++k) {
      const int c8g = c8g0 + RS * k;
      const int c = gc * MG * 8 + c8g;
      if (ok && c < nc) zr[(long long)c * P] = outs[(c8g * 16 + t) * OLD + ip];
    }
  }
}

// (terms, channel groups of 8, prefetch) per wave of each band-tile radial variant; variant
// 0 is the one-term, one-group kernel; variants with kRadPosMG > 0 are the position form
// with that many channel groups per wave (only with separable terms split off, which fixes
// the set of terms its tiles hold)
constexpr int kRadTG[kFhamRadialVariants] = {1, 2, 3, 3, 2, 3, 3, 1, 1, 1, 1, 1, 1, 1, 1};
constexpr int kRadGG[kFhamRadialVariants] = {1, 2, 2, 3, 2, 2, 3, 1, 1, 1, 1, 1, 1, 1, 1};
constexpr bool kRadPipe[kFhamRadialVariants] = {false, false, false, false, true, true, true,
                                                false, false, false, false, false, false, false, false};
constexpr int kRadPosMG[kFhamRadialVariants] = {0, 0, 0, 0, 0, 0, 0, 6, 3, 2, 3, 2, 2, 1, 2};
// position variants with the window in LDS (fham_radial_win_kernel): positions per workgroup
// and waves per workgroup (eight waves of two positions each measured no faster at helium:
// more waves per SIMD do not shorten it)
constexpr int kRadWinTW[kFhamRadialVariants] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 16, 16, 8, 16, 16};
constexpr int kRadWinNW[kFhamRadialVariants] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 4, 4, 4};
// the swizzled (unpadded) window, FP32 only: 54 KB at helium, three workgroups per CU
constexpr bool kRadWinSW[kFhamRadialVariants] = {false, false, false, false, false, false, false, false,
                                                 false, false, false, false, false, false, true};
// angular variants: waves splitting K, and whether one workgroup covers all row blocks of
// its position tile (reading each Z element once) or one row block
constexpr int kAngWaveChoices[kFhamAngularVariants] = {16, 8, 4, 8, 16, 4, 4};
constexpr bool kAngAllRows[kFhamAngularVariants] = {false, false, false, true, true, false, false};
// sparse variants: positions per lane (0: a GEMM form above)
constexpr int kAngSparsePV[kFhamAngularVariants] = {0, 0, 0, 0, 0, 1, 2};
constexpr int kAngMaxRowBlocks = 4;  // all-row variants: at most 64 channels

template <class T, int B>
static void launch_radial(int variant, int nc, int ncp, int n, int nterms, int ngroups, int tiles, const T* Xt,
                          const T* BT, const T* RP, T* Z, const int* differs, hipStream_t s) {
  if (kRadWinTW[variant] > 0) {
    if (!RP) throw Error(PP_ERR_INVALID, "position-form radial variant without position tiles");
    const int MG = kRadPosMG[variant], TW = kRadWinTW[variant], NW = kRadWinNW[variant];
    const bool SWv = kRadWinSW[variant];
    if (SWv && sizeof(T) != 4) throw Error(PP_ERR_INVALID, "the swizzled radial window is FP32 only");
    const int ntiles = (n + TW - 1) / TW;
    const long long nblocks = (long long)((nterms + 15) / 16) * n * ntiles * ((ngroups + MG - 1) / MG);
    const dim3 grid((unsigned)((nblocks + 7) / 8 * 8));
#define PP_RADW(G, W, V, S)                                                                                       \
  if constexpr (!S || sizeof(T) == 4) {                                                                           \
    if (MG == G && TW == W && NW == V && SWv == S) {                                                              \
      constexpr size_t lds = win_lds_bytes<T, B, G, W, V, S>();                                                   \
      if (lds > 160 * 1024) throw Error(PP_ERR_INVALID, "radial window does not fit in LDS");                     \
      PP_HIP(hipFuncSetAttribute((const void*)fham_radial_win_kernel<T, B, G, W, V, S>,                           \
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                          \
      hipLaunchKernelGGL((fham_radial_win_kernel<T, B, G, W, V, S>), grid, dim3(64 * V), lds, s, nc, n, nterms,   \
                         ngroups, ntiles, nblocks, ncp, Xt, RP, Z, differs);                                      \
      PP_HIP(hipGetLastError());                                                                                  \
      return;                                                                                                     \
    }                                                                                                             \
  }
    PP_RADW(3, 16, 4, false) PP_RADW(2, 16, 4, false) PP_RADW(2, 8, 4, false) PP_RADW(1, 16, 4, false)
    PP_RADW(2, 16, 4, true)
#undef PP_RADW
    throw Error(PP_ERR_INVALID, "radial window shape not instantiated");
  }
  if (kRadPosMG[variant] > 0) {
    const int MG = kRadPosMG[variant];
    if (!RP) throw Error(PP_ERR_INVALID, "position-form radial variant without position tiles");
    const long long nwaves = (long long)((nterms + 15) / 16) * n * n * ((ngroups + MG - 1) / MG);
    const long long nblocks = (nwaves + 3) / 4;
    const dim3 grid((unsigned)((nblocks + 7) / 8 * 8));
    if (MG == 6)
      hipLaunchKernelGGL((fham_radial_pos_kernel<T, B, 6>), grid, dim3(256), 0, s, nc, ncp, n, nterms, ngroups,
                         nwaves, Xt, RP, Z, differs);
    else if (MG == 3)
      hipLaunchKernelGGL((fham_radial_pos_kernel<T, B, 3>), grid, dim3(256), 0, s, nc, ncp, n, nterms, ngroups,
                         nwaves, Xt, RP, Z, differs);
    else
      hipLaunchKernelGGL((fham_radial_pos_kernel<T, B, 2>), grid, dim3(256), 0, s, nc, ncp, n, nterms, ngroups,
                         nwaves, Xt, RP, Z, differs);
    PP_HIP(hipGetLastError());
    return;
  }
  const int TG = kRadTG[variant], GG = kRadGG[variant];
  const long long nblocks =
      (long long)n * tiles * ((nterms + TG - 1) / TG) * ((ngroups + GG - 1) / GG);
  const dim3 grid((unsigned)((nblocks + 7) / 8 * 8));
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((fham_radial_kernel<T, B>), grid, dim3(64), 0, s, nc, ncp, n, nterms, ngroups, tiles,
                         nblocks, Xt, BT, Z, differs);
      break;
#define PP_RADV(V)                                                                                                   \
  case V:                                                                                                            \
    hipLaunchKernelGGL((fham_radial_reuse_kernel<T, B, kRadTG[V], kRadGG[V], kRadPipe[V]>), grid, dim3(64), 0, s,   \
                       nc, ncp, n, nterms, ngroups, tiles, nblocks, Xt, BT, Z, differs);                             \
    break;
      PP_RADV(1) PP_RADV(2) PP_RADV(3) PP_RADV(4) PP_RADV(5) PP_RADV(6)
#undef PP_RADV
    default:
      throw Error(PP_ERR_INVALID, "radial variant out of range");
  }
}

// Stage 2.  Workgroup of W waves per (16 rows c', 16 complex positions p); the waves
// split K = T nc and reduce through LDS.  A: [Mp][K] (row c', column t*nc + c).  The grid is
// one-dimensional over nrb row blocks x position tiles, renumbered so that the row blocks of
// one position tile run on one XCD together and read their Z columns through its L2.
constexpr int kAngBatch = 12;
template <class T, int W>
__global__ __launch_bounds__(64 * W) void fham_angular_kernel(int nc, int nrb, int K, int Kreal, long long P,
                                                                      const T* __restrict__ A,
                                                                      const int* __restrict__ tk,
                                                                      const typename MfmaT<T>::C2* __restrict__ Z,
                                                                      FScales<T> sc,
                                                                      const typename MfmaT<T>::C2* addend,
                                                                      typename MfmaT<T>::C2 alpha,
                                                                      const typename MfmaT<T>::C2* __restrict__ zs,
                                                                      typename MfmaT<T>::C2 beta,
                                                                      typename MfmaT<T>::C2* out,
                                                                      typename MfmaT<T>::C2* out2,
                                                                      typename MfmaT<T>::C2 alpha2,
                                                                      typename MfmaT<T>::C2 beta2,
                                                                      int* __restrict__ reset) {
  typedef typename MfmaT<T>::V4 V4;
  typedef typename MfmaT<T>::C2 C2;
  __shared__ V4 red[W - 1][2][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const long long nblocks = (long long)nrb * ((P + 15) / 16);
  const long long per_xcd = (nblocks + 7) / 8;
  const long long L = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (L >= nblocks) return;
  // the radial stage of this multiply has read the comparison flag: re-arm it
  if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
  const int m0 = (int)(L % nrb) * 16;
  const long long p0 = (L / nrb) * 16;
  const int nks = K >> 2;
  const int per = (nks + W - 1) / W;
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
      a[q] = ok ? A[((long long)(m0 >> 4) * nks + ks) * 64 + lane] * sc.s[tk[k]] : T(0);
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
  for (int q = 0; q < W - 1; ++q) {
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
    C2 v;
    v.x = 0;
    v.y = 0;
    if (zs) v = zs[idx];  // overlap term: its angular part is the identity
    o.x += beta.x * v.x - beta.y * v.y;
    o.y += beta.x * v.y + beta.y * v.x;
    out[idx] = o;
    if (out2) {  // second output from the same products (no addend)
      C2 o2;
      o2.x = alpha2.x * vr - alpha2.y * vi + beta2.x * v.x - beta2.y * v.y;
      o2.y = alpha2.x * vi + alpha2.y * vr + beta2.x * v.y + beta2.y * v.x;
      out2[idx] = o2;
    }
  }
}

// Stage 2, all row blocks per workgroup: a workgroup of W waves owns one tile of 16
// positions and every row block (MR of them), so each Z element is read once and feeds MR
// MFMAs; the waves split K and reduce through LDS in a tree (log2 W steps, W/2 slots).
constexpr int kAngBatchRows = 4;
template <class T, int W, int MR>
__global__ __launch_bounds__(64 * W) void fham_angular_rows_kernel(int nc, int K, int Kreal, long long P,
                                                                   const T* __restrict__ A,
                                                                   const int* __restrict__ tk,
                                                                   const typename MfmaT<T>::C2* __restrict__ Z,
                                                                   FScales<T> sc,
                                                                   const typename MfmaT<T>::C2* addend,
                                                                   typename MfmaT<T>::C2 alpha,
                                                                   const typename MfmaT<T>::C2* __restrict__ zs,
                                                                   typename MfmaT<T>::C2 beta,
                                                                   typename MfmaT<T>::C2* out,
                                                                   typename MfmaT<T>::C2* out2,
                                                                   typename MfmaT<T>::C2 alpha2,
                                                                   typename MfmaT<T>::C2 beta2,
                                                                   int* __restrict__ reset) {
  typedef typename MfmaT<T>::V4 V4;
  typedef typename MfmaT<T>::C2 C2;
  extern __shared__ char ang_red_raw[];
  V4* red = reinterpret_cast<V4*>(ang_red_raw);  // [W/2][MR][2][64]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
  const long long p0 = (long long)blockIdx.x * 16;
  const int nks = K >> 2;
  const int per = (nks + W - 1) / W;
  const int ks0 = wave * per, ks1 = min(nks, ks0 + per);
  const long long pc = p0 + lr;
  const bool col_ok = pc < P;
  V4 are[MR], aim[MR];
#pragma unroll
  for (int r = 0; r < MR; ++r) {
    are[r] = V4{0, 0, 0, 0};
    aim[r] = V4{0, 0, 0, 0};
  }
  for (int kb = ks0; kb < ks1; kb += kAngBatchRows) {
    T a[MR][kAngBatchRows];
    C2 z[kAngBatchRows];
#pragma unroll
    for (int q = 0; q < kAngBatchRows; ++q) {
      const int ks = kb + q;
      const int k = ks * 4 + lk;
      const bool ok = ks < ks1 && k < Kreal;
      const T f = ok ? sc.s[tk[k]] : T(0);
#pragma unroll
      for (int r = 0; r < MR; ++r) a[r][q] = ok ? A[((long long)r * nks + ks) * 64 + lane] * f : T(0);
      if (ok && col_ok) z[q] = Z[(long long)k * P + pc];
      else {
        z[q].x = 0;
        z[q].y = 0;
      }
    }
#pragma unroll
    for (int q = 0; q < kAngBatchRows; ++q)
#pragma unroll
      for (int r = 0; r < MR; ++r) {
        are[r] = MfmaT<T>::mma(a[r][q], z[q].x, are[r]);
        aim[r] = MfmaT<T>::mma(a[r][q], z[q].y, aim[r]);
      }
  }
  // tree: waves [h, 2h) hand their sums to waves [0, h)
#pragma unroll
  for (int h = W / 2; h >= 1; h /= 2) {
    if (wave >= h && wave < 2 * h) {
#pragma unroll
      for (int r = 0; r < MR; ++r) {
        red[(((wave - h) * MR + r) * 2 + 0) * 64 + lane] = are[r];
        red[(((wave - h) * MR + r) * 2 + 1) * 64 + lane] = aim[r];
      }
    }
    __syncthreads();
    if (wave < h) {
#pragma unroll
      for (int r = 0; r < MR; ++r) {
        are[r] += red[((wave * MR + r) * 2 + 0) * 64 + lane];
        aim[r] += red[((wave * MR + r) * 2 + 1) * 64 + lane];
      }
    }
    __syncthreads();
  }
  if (wave != 0 || !col_ok) return;
#pragma unroll
  for (int rb = 0; rb < MR; ++rb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = rb * 16 + (MfmaT<T>::kRowMajorOut ? lk * 4 + r : r * 4 + lk);
      if (c >= nc) continue;
      const long long idx = (long long)c * P + pc;
      C2 o;
      if (addend) o = addend[idx];
      else {
        o.x = 0;
        o.y = 0;
      }
      const T vr = are[rb][r], vi = aim[rb][r];
      o.x += alpha.x * vr - alpha.y * vi;
      o.y += alpha.x * vi + alpha.y * vr;
      C2 v;
      v.x = 0;
      v.y = 0;
      if (zs) v = zs[idx];  // overlap term: its angular part is the identity
      o.x += beta.x * v.x - beta.y * v.y;
      o.y += beta.x * v.y + beta.y * v.x;
      out[idx] = o;
      if (out2) {
        C2 o2;
        o2.x = alpha2.x * vr - alpha2.y * vi + beta2.x * v.x - beta2.y * v.y;
        o2.y = alpha2.x * vi + alpha2.y * vr + beta2.x * v.y + beta2.y * v.x;
        out2[idx] = o2;
      }
    }
  }
}

// Stage 2, sparse form.  The angular matrices are sparse: at helium 1269 of the
// 18 x 42 x 42 entries are non-zero (the H0 pieces are diagonal in the channel, each
// multipole and the dipole couple a few channels), so the GEMM above issues ~25 times the
// useful products.  Here Y[c'][p] = sum over the non-zeros (t, c) of row c' of
// f_t A_t[c'][c] Z[t][c][p] (f_t by value, or with scd from a per-potential array in device
// memory, so that a captured graph reads the scalings of each replay).  One workgroup of
// kAngSparseWaves waves per (output channel,
// chunk of 64 PV positions): the row's list -- Z row index t nc + c, coefficient, term --
// is split into contiguous parts, one per wave, read through scalar loads (uniform in the
// wave); lane l accumulates positions chunk + 64 r + l (r < PV), so every Z load is one
// contiguous 64-element read, and a wave issues all loads of its part (at most
// kAngSparseBatch entries per round) before it waits.  The parts are summed through LDS in
// wave order (deterministic).  Bandwidth-bound: each Z row is read by the workgroups of the
// output channels that use it; the workgroups of one position chunk run on one XCD
// (renumbered grid) and share its L2.
constexpr int kAngSparseWaves = 4;
constexpr int kAngSparseBatch = 12;  // list entries whose loads a wave issues together
template <class T, int PV>
__global__ __launch_bounds__(64 * kAngSparseWaves) void fham_angular_sparse_kernel(
    int nc, long long P, int nchunks, const int* __restrict__ rowptr, const int* __restrict__ colk,
    const T* __restrict__ coef, const int* __restrict__ cterm, const typename MfmaT<T>::C2* __restrict__ Z,
    FScales<T> sc, const double* __restrict__ scd, const int* __restrict__ cpot,
    const typename MfmaT<T>::C2* addend, typename MfmaT<T>::C2 alpha,
    const typename MfmaT<T>::C2* __restrict__ zs, typename MfmaT<T>::C2 beta, typename MfmaT<T>::C2* out,
    typename MfmaT<T>::C2* out2, typename MfmaT<T>::C2 alpha2, typename MfmaT<T>::C2 beta2, int* __restrict__ reset,
    double* __restrict__ npart) {
  typedef typename MfmaT<T>::C2 C2;
  __shared__ T part_re[kAngSparseWaves - 1][PV][64], part_im[kAngSparseWaves - 1][PV][64];
  const long long nblocks = (long long)nc * nchunks;
  const long long per_xcd = (nblocks + 7) / 8;
  const long long L = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (L >= nblocks) return;  // whole workgroups
  if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
  const int lane = threadIdx.x & 63;
  // wave-uniform indices (readfirstlane: the row list is then read through scalar loads)
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int chunk = (int)(L / nc), row = (int)(L - (L / nc) * nc);
  const long long p0 = (long long)chunk * 64 * PV + lane;
  const int r0 = rowptr[row], r1 = rowptr[row + 1];
  const int per = (r1 - r0 + kAngSparseWaves - 1) / kAngSparseWaves;
  const int j0 = min(r1, r0 + wave * per), j1 = min(r1, j0 + per);
  T re[PV], im[PV];
  bool ok[PV];
#pragma unroll
  for (int r = 0; r < PV; ++r) {
    re[r] = 0;
    im[r] = 0;
    ok[r] = p0 + 64 * r < P;
  }
  for (int jb = j0; jb < j1; jb += kAngSparseBatch) {
    C2 z[kAngSparseBatch][PV];
    T a[kAngSparseBatch];
#pragma unroll
    for (int q = 0; q < kAngSparseBatch; ++q) {
      const bool have = jb + q < j1;     // uniform
      const int j = have ? jb + q : jb;  // past the part's end: an in-bounds entry, weight 0
      const long long base = (long long)colk[j] * P;
      // the time scaling: by value per term, or (scd) per potential from device memory
      const T f = scd ? (T)scd[cpot[j]] : sc.s[cterm[j]];
      a[q] = have ? coef[j] * f : T(0);
#pragma unroll
      for (int r = 0; r < PV; ++r) {
        if (have && ok[r]) z[q][r] = Z[base + p0 + 64 * r];
        else {
          z[q][r].x = 0;
          z[q][r].y = 0;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kAngSparseBatch; ++q)
#pragma unroll
      for (int r = 0; r < PV; ++r) {
        re[r] = fma(a[q], z[q][r].x, re[r]);
        im[r] = fma(a[q], z[q][r].y, im[r]);
      }
  }
  if (wave > 0) {
#pragma unroll
    for (int r = 0; r < PV; ++r) {
      part_re[wave - 1][r][lane] = re[r];
      part_im[wave - 1][r][lane] = im[r];
    }
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int q = 0; q < kAngSparseWaves - 1; ++q)
#pragma unroll
    for (int r = 0; r < PV; ++r) {
      re[r] += part_re[q][r][lane];
      im[r] += part_im[q][r][lane];
    }
  double n1 = 0, n2 = 0;  // squared norms of the block's outputs (npart)
#pragma unroll
  for (int r = 0; r < PV; ++r) {
    if (!ok[r]) continue;
    const long long idx = (long long)row * P + p0 + 64 * r;
    C2 o;
    if (addend) o = addend[idx];
    else {
      o.x = 0;
      o.y = 0;
    }
    const T vr = re[r], vi = im[r];
    o.x += alpha.x * vr - alpha.y * vi;
    o.y += alpha.x * vi + alpha.y * vr;
    C2 v;
    v.x = 0;
    v.y = 0;
    if (zs) v = zs[idx];  // overlap term: its angular part is the identity
    o.x += beta.x * v.x - beta.y * v.y;
    o.y += beta.x * v.y + beta.y * v.x;
    out[idx] = o;
    n1 += (double)o.x * o.x + (double)o.y * o.y;
    if (out2) {
      C2 o2;
      o2.x = alpha2.x * vr - alpha2.y * vi + beta2.x * v.x - beta2.y * v.y;
      o2.y = alpha2.x * vi + alpha2.y * vr + beta2.x * v.y + beta2.y * v.x;
      out2[idx] = o2;
      n2 += (double)o2.x * o2.x + (double)o2.y * o2.y;
    }
  }
  if (npart) {  // [L][0] = |out|^2, [L][1] = |out2|^2 over the block, a shuffle tree in lane order
    for (int off = 32; off > 0; off >>= 1) {
      n1 += __shfl_down(n1, off, 64);
      n2 += __shfl_down(n2, off, 64);
    }
    if (lane == 0) {
      npart[2 * L] = n1;
      npart[2 * L + 1] = n2;
    }
  }
}

template <class T>
struct FOut {
  typename MfmaT<T>::C2* dst = nullptr;
  const typename MfmaT<T>::C2* addend = nullptr;
  double2 alpha = make_double2(1.0, 0.0), beta = make_double2(0.0, 0.0);
};

template <class T>
static typename MfmaT<T>::C2 to_c2(double2 v) {
  typename MfmaT<T>::C2 c;
  c.x = (T)v.x;
  c.y = (T)v.y;
  return c;
}

// Radial products of the separable terms R = a (x) b: Z_c = a X_c b^T for one (channel,
// term) per workgroup, X_c and the r2 pass held in LDS.  a, b: band rows [n][2 BW + 1]
// (a(i, i - BW + jj)).  Term j writes product slot first + j.  Skipped, like the band
// kernels, when the input equals the one whose products are held.
// Two optional jobs of the FP32 inner operator ride along:
//  * with cu1 (the upper band factors U1, U2 of the two radial overlaps, [j][BW + i - j],
//    band BW), workgroup y = 0 forms (S1 (x) S2) X_c in Cholesky form, U1^T U1 X_c U2^T U2,
//    by four triangular passes into slot cslot (the terms then take y = 1 ..): in FP32,
//    S (x) S applied entry by entry loses its small eigenvalues, this form keeps them;
//  * with Xt, the workgroups of channel c write its block, a slice each, into the
//    channel-minor copy Xt[j1][j2][c][part] that the radial kernels read (the transpose
//    launch is then skipped).
// A pass out[o] = sum_t coef(o, t) in[o + off + t] runs along one radial axis: wave k
// computes outputs 4k .. 4k + 3 (so the coefficients are wave-uniform: broadcast LDS reads),
// lane p the index across the axis, holding the 3 + taps inputs it needs in registers (4 +
// taps - 1 LDS reads for 4 outputs).  LDS rows have stride kSepLD, so that lanes reading down
// a column hit distinct banks.  The job's two coefficient arrays are copied into LDS with the
// channel's block, so no pass waits on a global load.
constexpr int kSepLD = 65;
constexpr int kSepR = 4;  // outputs per lane along a pass (16 waves x 4 >= 64)
template <class T, int TAPS, bool ALONG_R1, class CoefFn>
__device__ inline void sep_pass(const typename MfmaT<T>::C2* src, typename MfmaT<T>::C2* dst, int dst_ld, int n,
                                int off, int o0, int p, CoefFn coef) {
  typedef typename MfmaT<T>::C2 C2;
  C2 win[kSepR + TAPS - 1];
#pragma unroll
  for (int k = 0; k < kSepR + TAPS - 1; ++k) {
    const int i = o0 + off + k;
    if (i >= 0 && i < n && p < n) win[k] = ALONG_R1 ? src[i * kSepLD + p] : src[p * kSepLD + i];
    else {
      win[k].x = 0;
      win[k].y = 0;
    }
  }
#pragma unroll
  for (int r = 0; r < kSepR; ++r) {
    const int o = o0 + r;
    const int oc = min(o, n - 1);  // coefficient row (outputs past n are not stored)
    T re = 0, im = 0;
#pragma unroll
    for (int t = 0; t < TAPS; ++t) {
      const T a = coef(oc, t);
      re = fma(a, win[r + t].x, re);
      im = fma(a, win[r + t].y, im);
    }
    if (o < n && p < n) {
      C2 v;
      v.x = re;
      v.y = im;
      if (ALONG_R1) dst[o * dst_ld + p] = v;
      else dst[p * dst_ld + o] = v;
    }
  }
}

template <class T, int BW>
__global__ __launch_bounds__(1024) void fham_separable_kernel(int nc, int n, int first,
                                                              const T* __restrict__ SA, const T* __restrict__ SB,
                                                              const typename MfmaT<T>::C2* __restrict__ X,
                                                              typename MfmaT<T>::C2* __restrict__ Z,
                                                              const int* __restrict__ differs,
                                                              const T* __restrict__ cu1, const T* __restrict__ cu2,
                                                              int cslot, T* __restrict__ Xt, int ncp) {
  typedef typename MfmaT<T>::C2 C2;
  constexpr int W = 2 * BW + 1;
  if (differs && *differs == 0) return;  // input equals the one whose products Z holds
  extern __shared__ char sep_raw[];
  const int m = n * n;
  C2* Xs = reinterpret_cast<C2*>(sep_raw);  // n rows of kSepLD
  C2* Ts = Xs + n * kSepLD;
  T* Cs = reinterpret_cast<T*>(Ts + n * kSepLD);  // the job's coefficients: two arrays of cn values
  const int c = blockIdx.x;
  // with cu1, y = 0 is the Cholesky-form overlap job (dispatched first: its four dependent
  // passes are the longest chain) and y - 1 the separable term; else y is the term
  const bool chol = cu1 && blockIdx.y == 0;
  const int j = cu1 ? (int)blockIdx.y - 1 : (int)blockIdx.y;
  const C2* x = X + (size_t)c * m;
  const T* ca = chol ? cu1 : SA + (size_t)max(j, 0) * n * W;  // U1 or a
  const T* cb = chol ? cu2 : SB + (size_t)max(j, 0) * n * W;  // U2 or b
  const int cn = chol ? n * (BW + 1) : n * W;
  // staging: every load of a thread is issued before the first LDS store
  constexpr int kStageX = 4;                            // m <= 4096 (separable_fits)
  constexpr int kStageC = (2 * 64 * W + 1023) / 1024;   // 2 cn <= 2 n W, n <= 64
  C2 xv[kStageX];
  T cv[kStageC];
#pragma unroll
  for (int q = 0; q < kStageX; ++q) {  // past m: element 0 again (not stored)
    const int e = threadIdx.x + q * 1024;
    xv[q] = x[e < m ? e : 0];
  }
#pragma unroll
  for (int q = 0; q < kStageC; ++q) {  // past 2 cn: the first value again (not stored)
    const int e = threadIdx.x + q * 1024;
    const int ec = e < 2 * cn ? e : 0;
    cv[q] = ec < cn ? ca[ec] : cb[ec - cn];
  }
#pragma unroll
  for (int q = 0; q < kStageX; ++q) {
    const int e = threadIdx.x + q * 1024;
    if (e < m) {
      const int i = e / n;
      Xs[i * kSepLD + (e - i * n)] = xv[q];
    }
  }
#pragma unroll
  for (int q = 0; q < kStageC; ++q) {
    const int e = threadIdx.x + q * 1024;
    if (e < 2 * cn) Cs[e] = cv[q];
  }
  __syncthreads();
  if (Xt) {  // the channel's workgroups share the channel-minor copy, a slice each
    const int jobs = gridDim.y, per = (m + jobs - 1) / jobs;
    const int e0 = (int)blockIdx.y * per, e1 = min(m, e0 + per);
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
      const int i = e / n;
      reinterpret_cast<C2*>(Xt)[(size_t)e * ncp + c] = Xs[i * kSepLD + (e - i * n)];
    }
  }
  const int p = threadIdx.x & 63;
  const int o0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * kSepR;
  if (chol) {
    // U along an axis: out[o] = sum_t U(o, o + t) in[o + t] (U(o, o + t) past the last row
    // multiplies a zero input: its index is clamped into the array); U^T: out[o] =
    // sum_t U(o - BW + t, o) in[o - BW + t]
    auto Uc = [&](const T* u) {
      const int nn = n;
      return [u, nn](int o, int t) { return u[min(o + t, nn - 1) * (BW + 1) + BW - t]; };
    };
    auto UT = [&](const T* u) { return [u](int o, int t) { return u[o * (BW + 1) + t]; }; };
    const T* u1 = Cs;
    const T* u2 = Cs + cn;
    sep_pass<T, BW + 1, false>(Xs, Ts, kSepLD, n, 0, o0, p, Uc(u2));    // U2 along r2
    __syncthreads();
    sep_pass<T, BW + 1, false>(Ts, Xs, kSepLD, n, -BW, o0, p, UT(u2));  // U2^T along r2
    __syncthreads();
    sep_pass<T, BW + 1, true>(Xs, Ts, kSepLD, n, 0, o0, p, Uc(u1));     // U1 along r1
    __syncthreads();
    C2* z = Z + ((size_t)cslot * nc + c) * m;
    sep_pass<T, BW + 1, true>(Ts, z, n, n, -BW, o0, p, UT(u1));         // U1^T along r1, to the slot
    return;
  }
  const T* fa = Cs;
  const T* fb = Cs + cn;
  // T = X b^T along r2: T[i1][o] = sum_d b[o][d] X[i1][o - BW + d]
  sep_pass<T, W, false>(Xs, Ts, kSepLD, n, -BW, o0, p, [fb](int o, int t) { return fb[o * W + t]; });
  __syncthreads();
  // Z = a T along r1: Z[o][i2] = sum_jj a[o][jj] T[o - BW + jj][i2]
  C2* z = Z + ((size_t)(first + j) * nc + c) * m;
  sep_pass<T, W, true>(Ts, z, n, n, -BW, o0, p, [fa](int o, int t) { return fa[o * W + t]; });
}

template <class T>
__global__ void fham_fill_zero_kernel(long long count, T* __restrict__ p) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    p[i] = T(0);
}

// o1.dst = o1.addend + o1.alpha Y + o1.beta Zs, o2 likewise without addend (may be NULL);
// Y = sum_t f_t A_t Z_t, Zs = the overlap product.  ref / differs: reuse of the radial
// products when src equals ref (FP64 only).
template <class T>
static void fham_run(const FHam& F, const double* scalings, const typename MfmaT<T>::C2* src, const FOut<T>& o1,
                     const FOut<T>* o2, const DeviceBuffer<T>& Ab, const DeviceBuffer<T>& Db,
                     const DeviceBuffer<T>& RPb, const DeviceBuffer<T>& SAb, const DeviceBuffer<T>& SBb, DeviceBuffer<T>& Zb,
                     DeviceBuffer<T>& Xb, const T* ref, int* differs, hipStream_t s, const double* sc_dev,
                     const typename MfmaT<T>::C2* zs_ext = nullptr, double* npart = nullptr) {
  const int nc = F.nc, n = F.n, nterms = F.nterms;
  const long long P = (long long)n * n;
  auto nz = [](double2 v) { return v.x != 0.0 || v.y != 0.0; };
  const bool with_s = nz(o1.beta) || (o2 && nz(o2->beta));
  // the overlap product is formed here unless zs_ext brings it
  const bool own_s = with_s && !zs_ext;
  // FP32 with the Cholesky factors: the overlap product by the four triangular passes
  const bool chol = sizeof(T) == 4 && own_s && F.csd == F.bw && F.sep_enabled;
  PP_REQUIRE(!own_s || F.has_overlap || chol, "overlap coefficient given but the operator has no overlap term");
  PP_REQUIRE(!(o2 && o2->addend), "the second output takes no addend");
  const int nrad = nterms + (own_s ? 1 : 0);  // radial terms formed (the overlap term is last)
  const size_t zneed = (size_t)(nterms + ((F.has_overlap || chol) ? 1 : 0)) * nc * P * 2;
  if (Zb.count < zneed) Zb.resize(zneed);
  const int ngroups = (nc + 7) / 8;
  const int ncp = ngroups * 8;
  const size_t xneed = (size_t)P * ncp * 2;
  if (Xb.count < xneed) {
    Xb.resize(xneed);
    // padding channels stay zero (the separable kernel writes only the nc real ones)
    hipLaunchKernelGGL(fham_fill_zero_kernel<T>, dim3((unsigned)std::min<long long>(4096, ceil_div((long long)xneed, 256LL))),
                       dim3(256), 0, s, (long long)xneed, Xb.ptr);
  }
  // separable terms (and, in FP32, the Cholesky-form overlap) run first when their kernel
  // can also write the channel-minor copy: no transpose launch.  Not with a reuse reference
  // (FP64): the transpose kernel then compares the input with the held one, and the radial
  // and separable stages are skipped when it is unchanged
  const int nsep_terms = F.sep_enabled ? nterms - F.nband : 0;
  const int nsep_run = nsep_terms + (own_s && !chol && F.sep_enabled ? 1 : 0);
  const int sep_jobs = nsep_run + (chol ? 1 : 0);
  const bool fused_xt = F.sep_enabled && sep_jobs > 0 && !ref;
  auto launch_separable = [&](T* xt) {
    if (sep_jobs <= 0) return;
    typedef typename MfmaT<T>::C2 C2s;
    const size_t lds = 2 * (size_t)n * kSepLD * sizeof(C2s) + 2 * (size_t)n * (2 * F.bw + 1) * sizeof(T);
    const T* cu1 = chol ? reinterpret_cast<const T*>(F.cu1_32.ptr) : nullptr;
    const T* cu2 = chol ? reinterpret_cast<const T*>(F.cu2_32.ptr) : nullptr;
    switch (F.bw) {
#define PP_SEP(B)                                                                                                 \
  case B:                                                                                                         \
    PP_HIP(hipFuncSetAttribute((const void*)fham_separable_kernel<T, B>,                                          \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                            \
    hipLaunchKernelGGL((fham_separable_kernel<T, B>), dim3(nc, sep_jobs), dim3(1024), lds, s, nc, n, F.nband,      \
                       SAb.ptr, SBb.ptr, src, reinterpret_cast<C2s*>(Zb.ptr), differs, cu1, cu2, nterms, xt, ncp);  \
    break;
      PP_SEP(1) PP_SEP(2) PP_SEP(3) PP_SEP(4) PP_SEP(5) PP_SEP(6) PP_SEP(7) PP_SEP(8)
#undef PP_SEP
      default:
        throw Error(PP_ERR_INVALID, "factored multiply supports band half-widths 1..8");
    }
    PP_HIP(hipGetLastError());
  };
  if (fused_xt) launch_separable(Xb.ptr);
  else
    hipLaunchKernelGGL(fham_transpose_kernel<T>, dim3((unsigned)std::min<long long>(4096, ceil_div(P * ncp * 2, 256))),
                       dim3(256), 0, s, nc, ncp, n, reinterpret_cast<const T*>(src), Xb.ptr, ref, differs);
  const int tiles = (n + 15) / 16;
  const int variant = sizeof(T) == 8 ? F.rad_variant64 : F.rad_variant32;
  // band-tile terms first; with separable terms those (and the overlap term) follow
  const int nrad_band = F.sep_enabled ? F.nband : nrad;
  PP_REQUIRE(kRadPosMG[variant] == 0 || F.sep_enabled, "position-form radial variant needs the separable split");
  if (nrad_band > 0) {
    switch (F.bw) {
#define PP_RAD(B)                                                                                                    \
  case B:                                                                                                            \
    launch_radial<T, B>(variant, nc, ncp, n, nrad_band, ngroups, tiles, Xb.ptr, Db.ptr, RPb.ptr, Zb.ptr, differs,   \
                        s);                                                                                          \
    break;
      PP_RAD(1) PP_RAD(2) PP_RAD(3) PP_RAD(4) PP_RAD(5) PP_RAD(6) PP_RAD(7) PP_RAD(8)
#undef PP_RAD
      default:
        throw Error(PP_ERR_INVALID, "factored multiply supports band half-widths 1..8");
    }
  }
  if (!fused_xt && (F.sep_enabled || chol)) launch_separable(nullptr);
  FScales<T> sc;
  for (int t = 0; t < kMaxFTerms; ++t) sc.s[t] = t < nterms ? (T)scalings[F.term_pot[t]] : T(0);
  const int nrb = (nc + 15) / 16;
  const long long ablocks = (long long)nrb * ceil_div(P, 16);
  typedef typename MfmaT<T>::C2 C2;
  const C2* Zc = reinterpret_cast<const C2*>(Zb.ptr);
  const dim3 agrid((unsigned)((ablocks + 7) / 8 * 8));
  const C2* zsp = zs_ext ? zs_ext : with_s ? Zc + (size_t)nterms * nc * P : nullptr;  // the overlap product
  const int avar = sizeof(T) == 8 ? F.ang_variant64 : F.ang_variant32;
  const int waves = kAngWaveChoices[avar];
  PP_REQUIRE(!npart || kAngSparsePV[avar] > 0, "output norm partials need the sparse angular form");
  if (kAngSparsePV[avar] > 0) {
    const int PV = kAngSparsePV[avar];
    const int nchunks = (int)ceil_div(P, 64LL * PV);
    const long long nb = (long long)nc * nchunks;
    const dim3 sgrid((unsigned)((nb + 7) / 8 * 8));
    const T* coef = sizeof(T) == 8 ? reinterpret_cast<const T*>(F.sp_coef64.ptr)
                                   : reinterpret_cast<const T*>(F.sp_coef32.ptr);
#define PP_ANGS(V)                                                                                                   \
  hipLaunchKernelGGL((fham_angular_sparse_kernel<T, V>), sgrid, dim3(64 * kAngSparseWaves), 0, s, nc, P, nchunks,   \
                     F.sp_rowptr.ptr, F.sp_k.ptr, coef, F.sp_term.ptr, Zc, sc, sc_dev, F.sp_pot.ptr, o1.addend,     \
                     to_c2<T>(o1.alpha),                                                                             \
                     zsp, to_c2<T>(o1.beta), o1.dst,                     \
                     o2 ? o2->dst : nullptr, to_c2<T>(o2 ? o2->alpha : make_double2(0, 0)),                          \
                     to_c2<T>(o2 ? o2->beta : make_double2(0, 0)), differs, npart)
    if (PV == 1) PP_ANGS(1);
    else PP_ANGS(2);
#undef PP_ANGS
    PP_HIP(hipGetLastError());
    return;
  }
  if (kAngAllRows[avar] && nrb <= kAngMaxRowBlocks) {
    typedef typename MfmaT<T>::V4 V4;
    const dim3 rgrid((unsigned)ceil_div(P, 16));
    const size_t lds = (size_t)(waves / 2) * nrb * 2 * 64 * sizeof(V4);
#define PP_ANGR(W, MR)                                                                                           \
  if (waves == W && nrb == MR) {                                                                                 \
    PP_HIP(hipFuncSetAttribute((const void*)fham_angular_rows_kernel<T, W, MR>,                                  \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                           \
    hipLaunchKernelGGL((fham_angular_rows_kernel<T, W, MR>), rgrid, dim3(64 * W), lds, s, nc, F.Kp, nterms * nc, \
                       P, Ab.ptr, F.tk.ptr, Zc, sc, o1.addend, to_c2<T>(o1.alpha),                               \
                       zsp, to_c2<T>(o1.beta), o1.dst,              \
                       o2 ? o2->dst : nullptr, to_c2<T>(o2 ? o2->alpha : make_double2(0, 0)),                    \
                       to_c2<T>(o2 ? o2->beta : make_double2(0, 0)), differs);                                   \
  }
    PP_ANGR(8, 1) PP_ANGR(8, 2) PP_ANGR(8, 3) PP_ANGR(8, 4)
    PP_ANGR(16, 1) PP_ANGR(16, 2) PP_ANGR(16, 3) PP_ANGR(16, 4)
#undef PP_ANGR
    PP_HIP(hipGetLastError());
    return;
  }
#define PP_ANG(W)                                                                                                    \
  hipLaunchKernelGGL((fham_angular_kernel<T, W>), agrid, dim3(64 * W), 0, s, nc, nrb, F.Kp, nterms * nc, P, Ab.ptr, \
                     F.tk.ptr, Zc, sc, o1.addend, to_c2<T>(o1.alpha), zsp,       \
                     to_c2<T>(o1.beta), o1.dst, o2 ? o2->dst : nullptr,                                              \
                     to_c2<T>(o2 ? o2->alpha : make_double2(0, 0)), to_c2<T>(o2 ? o2->beta : make_double2(0, 0)),   \
                     differs)
  if (waves == 4) PP_ANG(4);
  else if (waves == 8) PP_ANG(8);
  else PP_ANG(16);
#undef PP_ANG
  PP_HIP(hipGetLastError());
}

void FHam::set_cholesky32(const std::vector<double>& u1, const std::vector<double>& u2, int sd) const {
  PP_REQUIRE(sd >= 0 && sd <= 2 * bw, "overlap band wider than the operator's");
  PP_REQUIRE(u1.size() == (size_t)n * (sd + 1) && u2.size() == (size_t)n * (sd + 1), "overlap factor size mismatch");
  std::vector<float> f1(u1.begin(), u1.end()), f2(u2.begin(), u2.end());
  cu1_32.upload(f1.data(), f1.size(), nullptr);
  cu2_32.upload(f2.data(), f2.size(), nullptr);
  PP_HIP(hipStreamSynchronize(nullptr));
  csd = sd;
}

void fham_apply64(const FHam& f, const double* scalings, const double2* src, double2* dst, const double2* addend,
                  double2 alpha, hipStream_t s) {
  f.apply64(scalings, src, dst, addend, alpha, s);
}

void FHam::apply64(const double* scalings, const double2* src, double2* dst, const double2* addend, double2 alpha,
                   hipStream_t s, double2 beta, const double* sc_dev) const {
  FHamOut o;
  o.dst = dst;
  o.addend = addend;
  o.alpha = alpha;
  o.beta = beta;
  apply64_ex(scalings, src, o, nullptr, nullptr, s, sc_dev);
}

void FHam::apply64_ex(const double* scalings, const double2* src, const FHamOut& o1, const FHamOut* o2,
                      const double2* reuse_ref, hipStream_t s, const double* sc_dev, double* npart) const {
  auto nz = [](double2 v) { return v.x != 0.0 || v.y != 0.0; };
  const int nrad = nterms + ((nz(o1.beta) || (o2 && nz(o2->beta))) ? 1 : 0);
  PP_REQUIRE(!reuse_ref || w64_terms >= nrad, "held radial products do not cover the requested terms");
  if (reuse_ref && reuse_flag.count == 0) {
    reuse_flag.resize(1);
    PP_HIP(hipMemsetAsync(reuse_flag.ptr, 0, sizeof(int), s));  // re-armed by every multiply that reads it
  }
  ++gen64;
  FOut<double> a1, a2;
  a1.dst = o1.dst;
  a1.addend = o1.addend;
  a1.alpha = o1.alpha;
  a1.beta = o1.beta;
  if (o2) {
    a2.dst = o2->dst;
    a2.addend = o2->addend;
    a2.alpha = o2->alpha;
    a2.beta = o2->beta;
  }
  fham_run<double>(*this, scalings, src, a1, o2 ? &a2 : nullptr, A64, bt64, rp64, sa64, sb64, W64, X64,
                   reinterpret_cast<const double*>(reuse_ref), reuse_ref ? reuse_flag.ptr : nullptr, s, sc_dev,
                   nullptr, npart);
  // when the products were reused they cover at least nrad terms; when recomputed, exactly nrad
  w64_terms = nrad;
}

void FHam::apply32(const double* scalings, const float2* src, float2* dst, const float2* addend, double2 alpha,
                   hipStream_t s, double2 beta, const double* sc_dev, const float2* zs_ext) const {
  FOut<float> o;
  o.dst = dst;
  o.addend = addend;
  o.alpha = alpha;
  o.beta = beta;
  fham_run<float>(*this, scalings, src, o, nullptr, A32, bt32, rp32, sa32, sb32, W32, X32, nullptr, nullptr, s,
                  sc_dev, zs_ext);
}

bool FHam::sparse_angular() const { return kAngSparsePV[ang_variant64] > 0 && kAngSparsePV[ang_variant32] > 0; }

long long FHam::norm_blocks64() const {
  const int pv = kAngSparsePV[ang_variant64];
  return pv > 0 ? (long long)nc * ceil_div((long long)n * n, 64LL * pv) : 0;
}

}  // namespace pp

using namespace pp;

// Band tiles in MFMA lane order [t][i1][jj][tile][ks][lane] from dia[t][i1][jj][d][i2]
// (element B[k][m] = dia[t][i1][jj][k - m][16 tile + m], zero outside the band).
static std::vector<double> band_tiles(int nterms, int n, int bw, const double* dia) {
  const int w = 2 * bw + 1;
  const int tiles = (n + 15) / 16;
  const int KS = (16 + 2 * bw + 3) / 4;
  std::vector<double> bt((size_t)nterms * n * w * tiles * KS * 64, 0.0);
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
  return bt;
}

// Position tiles [term tile][i1 n + i2][ks][lane] from dia[t][i1][jj][d][i2]: lane l of step
// ks holds R[k][t] = dia[t][i1][jj][d][i2] at k = 4 ks + l / 16 = jj w + d, t = 16 tile + l % 16.
static std::vector<double> position_tiles(int nterms, int n, int bw, const double* dia) {
  const int w = 2 * bw + 1;
  const int KP = (w * w + 3) / 4;
  const int ntn = (nterms + 15) / 16;
  std::vector<double> rp((size_t)ntn * n * n * KP * 64, 0.0);
  for (int t = 0; t < nterms; ++t)
    for (int i1 = 0; i1 < n; ++i1)
      for (int jj = 0; jj < w; ++jj)
        for (int d = 0; d < w; ++d) {
          const int k = jj * w + d;
          const int lane = (k % 4) * 16 + t % 16;
          for (int i2 = 0; i2 < n; ++i2)
            rp[((((size_t)(t / 16) * n + i1) * n + i2) * KP + k / 4) * 64 + lane] =
                dia[((((size_t)t * n + i1) * w + jj) * w + d) * n + i2];
        }
  return rp;
}

// A term whose band D[i1][jj][d][i2] equals a[i1][jj] b[i2][d] (entries within 1e-13 of the
// largest) is separable; a, b are read from the row and column through its largest entry.
static bool separable_factors(const double* D, int n, int bw, std::vector<double>& a, std::vector<double>& b) {
  const int w = 2 * bw + 1;
  const size_t cnt = (size_t)n * w * w * n;
  size_t best = 0;
  double mx = 0.0;
  for (size_t e = 0; e < cnt; ++e)
    if (std::abs(D[e]) > mx) {
      mx = std::abs(D[e]);
      best = e;
    }
  a.assign((size_t)n * w, 0.0);
  b.assign((size_t)n * w, 0.0);
  if (mx == 0.0) return true;
  const int i2s = (int)(best % n), ds = (int)((best / n) % w), jjs = (int)((best / ((size_t)n * w)) % w);
  const int i1s = (int)(best / ((size_t)n * w * w));
  const double piv = D[best];
  for (int i1 = 0; i1 < n; ++i1)
    for (int jj = 0; jj < w; ++jj) a[(size_t)i1 * w + jj] = D[(((size_t)i1 * w + jj) * w + ds) * n + i2s];
  for (int d = 0; d < w; ++d)
    for (int i2 = 0; i2 < n; ++i2) b[(size_t)i2 * w + d] = D[(((size_t)i1s * w + jjs) * w + d) * n + i2] / piv;
  for (int i1 = 0; i1 < n; ++i1)
    for (int jj = 0; jj < w; ++jj)
      for (int d = 0; d < w; ++d)
        for (int i2 = 0; i2 < n; ++i2) {
          const double v = D[(((size_t)i1 * w + jj) * w + d) * n + i2];
          if (std::abs(v - a[(size_t)i1 * w + jj] * b[(size_t)i2 * w + d]) > 1e-13 * mx) return false;
        }
  return true;
}

// the separable kernel holds X_c, one pass and both factors in LDS, and maps one thread to
// each of at most 64 columns
static bool separable_fits(int n) {  // FP64 block and pass in LDS, with two coefficient arrays (band <= 8)
  return n <= 64 && 2 * (size_t)n * kSepLD * sizeof(double2) + 2 * (size_t)n * 17 * sizeof(double) <= 160 * 1024;
}

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
    // separable terms go last, in their own kernel; the others keep band tiles
    const int w = 2 * bw + 1;
    const size_t dstride = (size_t)n * w * w * n;
    std::vector<std::vector<double>> fa(nterms), fb(nterms);
    std::vector<int> sep(nterms, 0), order;
    const bool sep_ok = separable_fits(n);
    for (int t = 0; t < nterms; ++t) sep[t] = sep_ok && separable_factors(dia + t * dstride, n, bw, fa[t], fb[t]);
    for (int t = 0; t < nterms; ++t)
      if (!sep[t]) order.push_back(t);
    F.nband = (int)order.size();
    for (int t = 0; t < nterms; ++t)
      if (sep[t]) order.push_back(t);
    F.sep_enabled = sep_ok;
    F.term_pot.resize(nterms);
    for (int k = 0; k < nterms; ++k) F.term_pot[k] = term_potential[order[k]];
    std::vector<double> dband((size_t)F.nband * dstride);
    for (int k = 0; k < F.nband; ++k) std::copy(dia + order[k] * dstride, dia + (order[k] + 1) * dstride,
                                                dband.begin() + (size_t)k * dstride);
    std::vector<double> sav, sbv;
    for (int k = F.nband; k < nterms; ++k) {
      sav.insert(sav.end(), fa[order[k]].begin(), fa[order[k]].end());
      sbv.insert(sbv.end(), fb[order[k]].begin(), fb[order[k]].end());
    }
    const int Mp = (nchan + 15) / 16 * 16;
    // A(c', t*nc + c) = A_t[c'][c], rows padded to a multiple of 16 with zeros, stored in
    // MFMA operand order: [row block][K step][lane], lane l holding A(16 rb + l % 16,
    // 4 ks + l / 16), so that a wave reads its 64 operands as one coalesced load
    const int nks = F.Kp / 4;
    std::vector<double> a((size_t)Mp * F.Kp, 0.0);
    for (int t = 0; t < nterms; ++t)
      for (int r = 0; r < nchan; ++r)
        for (int c = 0; c < nchan; ++c) {
          const int k = t * nchan + c;
          a[((size_t)(r / 16) * nks + k / 4) * 64 + (k % 4) * 16 + r % 16] =
              A[((size_t)order[t] * nchan + r) * nchan + c];
        }
    std::vector<int> tk(F.Kp);
    for (int k = 0; k < F.Kp; ++k) tk[k] = std::min(k / nchan, nterms - 1);  // term of column k
    // the same matrix as row lists of its non-zeros for the sparse angular form: row c' holds
    // (Z row t nc + c, A_t[c'][c], t) in increasing (t, c), terms in the reordered numbering
    std::vector<int> sp_rowptr(nchan + 1, 0), sp_k, sp_term, sp_pot;
    std::vector<double> sp_coef;
    for (int r = 0; r < nchan; ++r) {
      for (int t = 0; t < nterms; ++t)
        for (int c = 0; c < nchan; ++c) {
          const double v = A[((size_t)order[t] * nchan + r) * nchan + c];
          if (v == 0.0) continue;
          sp_k.push_back(t * nchan + c);
          sp_coef.push_back(v);
          sp_term.push_back(t);
          sp_pot.push_back(F.term_pot[t]);
        }
      sp_rowptr[r + 1] = (int)sp_k.size();
    }
    if (sp_k.empty()) {  // an all-zero operator: keep the device lists non-empty
      sp_k.push_back(0);
      sp_coef.push_back(0.0);
      sp_term.push_back(0);
      sp_pot.push_back(0);
    }
    std::vector<double> bt = band_tiles(F.nband, n, bw, dband.data());
    const size_t nbt = bt.size();
    std::vector<double> rp;
    if (F.sep_enabled && F.nband > 0) rp = position_tiles(F.nband, n, bw, dband.data());
    try {
      hipStream_t s = nullptr;
      F.A64.upload(a.data(), a.size(), s);
      F.tk.upload(tk.data(), tk.size(), s);
      F.sp_rowptr.upload(sp_rowptr.data(), sp_rowptr.size(), s);
      F.sp_k.upload(sp_k.data(), sp_k.size(), s);
      F.sp_term.upload(sp_term.data(), sp_term.size(), s);
      F.sp_pot.upload(sp_pot.data(), sp_pot.size(), s);
      F.sp_coef64.upload(sp_coef.data(), sp_coef.size(), s);
      std::vector<float> spf(sp_coef.begin(), sp_coef.end());
      F.sp_coef32.upload(spf.data(), spf.size(), s);
      std::vector<float> af(a.begin(), a.end());
      F.A32.upload(af.data(), af.size(), s);
      F.bt64.upload(bt.data(), nbt, s);
      std::vector<float> bf(bt.begin(), bt.end());
      F.bt32.upload(bf.data(), nbt, s);
      if (!rp.empty()) {
        F.rp64.upload(rp.data(), rp.size(), s);
        std::vector<float> rf(rp.begin(), rp.end());
        F.rp32.upload(rf.data(), rf.size(), s);
      }
      F.sa64.upload(sav.data(), sav.size(), s);
      F.sb64.upload(sbv.data(), sbv.size(), s);
      std::vector<float> saf(sav.begin(), sav.end()), sbf(sbv.begin(), sbv.end());
      F.sa32.upload(saf.data(), saf.size(), s);
      F.sb32.upload(sbf.data(), sbf.size(), s);
      PP_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

extern "C" int pp_fham_band_terms(const pp_fham* f, int* band_terms) {
  return guarded([&] {
    PP_REQUIRE(f && band_terms, "null argument");
    *band_terms = f->f.sep_enabled ? f->f.nband : f->f.nterms;
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

// Whether radial variant v can run for this operator in the given precision: the position
// forms need the separable split; the LDS-window forms a window that fits in 160 KiB (the
// swizzled window: FP32 only).
static bool radial_window_fits(const FHam& F, int v, int precision) {
  if (kRadWinTW[v] == 0) return true;
  if (kRadWinSW[v] && precision == 64) return false;  // the swizzled window is FP32 only
  return win_lds_bytes(F.bw, kRadPosMG[v], kRadWinTW[v], kRadWinNW[v], precision == 64 ? 8 : 4, kRadWinSW[v]) <=
         160 * 1024;
}

extern "C" int pp_fham_set_radial_variant(pp_fham* f, int precision, int variant) {
  return guarded([&] {
    PP_REQUIRE(f && (precision == 64 || precision == 32), "precision must be 64 or 32");
    PP_REQUIRE(variant >= 0 && variant < kFhamRadialVariants, "radial variant out of range");
    PP_REQUIRE(kRadPosMG[variant] == 0 || f->f.sep_enabled,
               "position-form radial variants need the separable split (n <= 64)");
    PP_REQUIRE(!(kRadWinSW[variant] && precision == 64), "the swizzled radial window is FP32 only");
    PP_REQUIRE(radial_window_fits(f->f, variant, precision), "radial window does not fit in LDS");
    (precision == 64 ? f->f.rad_variant64 : f->f.rad_variant32) = variant;
  });
}

extern "C" int pp_fham_get_radial_variant(const pp_fham* f, int precision, int* variant) {
  return guarded([&] {
    PP_REQUIRE(f && variant && (precision == 64 || precision == 32), "invalid argument");
    *variant = precision == 64 ? f->f.rad_variant64 : f->f.rad_variant32;
  });
}

extern "C" int pp_fham_set_angular_waves(pp_fham* f, int precision, int waves) {
  return guarded([&] {
    PP_REQUIRE(f && (precision == 64 || precision == 32), "precision must be 64 or 32");
    int v = -1;
    for (int i = 0; i < kFhamAngularVariants; ++i)
      if (!kAngAllRows[i] && !kAngSparsePV[i] && kAngWaveChoices[i] == waves) v = i;
    PP_REQUIRE(v >= 0, "angular wave count must be 16, 8 or 4");
    (precision == 64 ? f->f.ang_variant64 : f->f.ang_variant32) = v;
  });
}

extern "C" int pp_fham_set_angular_variant(pp_fham* f, int precision, int variant) {
  return guarded([&] {
    PP_REQUIRE(f && (precision == 64 || precision == 32), "precision must be 64 or 32");
    PP_REQUIRE(variant >= 0 && variant < kFhamAngularVariants, "angular variant out of range");
    (precision == 64 ? f->f.ang_variant64 : f->f.ang_variant32) = variant;
  });
}

extern "C" int pp_fham_get_angular_variant(const pp_fham* f, int precision, int* variant) {
  return guarded([&] {
    PP_REQUIRE(f && variant && (precision == 64 || precision == 32), "invalid argument");
    *variant = precision == 64 ? f->f.ang_variant64 : f->f.ang_variant32;
  });
}

extern "C" int pp_fham_get_angular_waves(const pp_fham* f, int precision, int* waves) {
  return guarded([&] {
    PP_REQUIRE(f && waves && (precision == 64 || precision == 32), "invalid argument");
    *waves = kAngWaveChoices[precision == 64 ? f->f.ang_variant64 : f->f.ang_variant32];
  });
}

// Times every radial variant on a zero input (one warm-up and `trials` timed multiplies each,
// HIP events on `stream`, in three interleaved rounds keeping each variant's fastest) and keeps
// the fastest for that precision; then the angular forms with it, the same way.
extern "C" int pp_fham_tune(pp_fham* f, int precision, int trials, int* best, double* timings_ms, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(f && (precision == 64 || precision == 32) && trials >= 1, "invalid argument");
    FHam& F = f->f;
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)F.nc * F.n * F.n;
    const size_t esz = precision == 64 ? sizeof(double2) : sizeof(float2);
    DeviceBuffer<char> src, dst;
    src.resize(N * esz);
    dst.resize(N * esz);
    PP_HIP(hipMemsetAsync(src.ptr, 0, N * esz, s));
    std::vector<double> sc(F.npot, 1.0);
    int& var = precision == 64 ? F.rad_variant64 : F.rad_variant32;
    auto run = [&] {
      if (precision == 64)
        F.apply64(sc.data(), reinterpret_cast<const double2*>(src.ptr), reinterpret_cast<double2*>(dst.ptr), nullptr,
                  make_double2(1.0, 0.0), s);
      else
        F.apply32(sc.data(), reinterpret_cast<const float2*>(src.ptr), reinterpret_cast<float2*>(dst.ptr), nullptr,
                  make_double2(1.0, 0.0), s);
    };
    hipEvent_t e0, e1;
    PP_HIP(hipEventCreate(&e0));
    PP_HIP(hipEventCreate(&e1));
    // ms per multiply of the current choice: one warm-up, then `trials` timed multiplies
    auto time_one = [&]() {
      run();
      PP_HIP(hipEventRecord(e0, s));
      for (int k = 0; k < trials; ++k) run();
      PP_HIP(hipEventRecord(e1, s));
      PP_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      PP_HIP(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms / trials;
    };
    // every candidate is timed in kRounds interleaved rounds and keeps its fastest, so that a
    // slow stretch of the device (clock ramp, another process) does not decide the choice
    constexpr int kRounds = 3;
    const int var0 = var;
    std::vector<double> rad(kFhamRadialVariants, 1e300);
    for (int round = 0; round < kRounds; ++round)
      for (int v = 0; v < kFhamRadialVariants; ++v) {
        if ((kRadPosMG[v] > 0 && !F.sep_enabled) || !radial_window_fits(F, v, precision)) continue;
        var = v;
        rad[v] = std::min(rad[v], time_one());
      }
    int best_v = -1;
    for (int v = 0; v < kFhamRadialVariants; ++v) {
      const bool avail = rad[v] < 1e300;
      if (timings_ms) timings_ms[v] = avail ? rad[v] : -1.0;  // -1: not available for this operator
      if (avail && (best_v < 0 || rad[v] < rad[best_v])) best_v = v;
    }
    if (best_v < 0) best_v = var0;
    var = best_v;
    // then the angular stage's form, with the chosen radial shape
    int& avar = precision == 64 ? F.ang_variant64 : F.ang_variant32;
    std::vector<double> ang(kFhamAngularVariants, 1e300);
    for (int round = 0; round < kRounds; ++round)
      for (int v = 0; v < kFhamAngularVariants; ++v) {
        avar = v;
        ang[v] = std::min(ang[v], time_one());
      }
    int best_av = 0;
    for (int v = 0; v < kFhamAngularVariants; ++v) {
      if (timings_ms) timings_ms[kFhamRadialVariants + v] = ang[v];
      if (ang[v] < ang[best_av]) best_av = v;
    }
    avar = best_av;
    PP_HIP(hipEventDestroy(e0));
    PP_HIP(hipEventDestroy(e1));
    if (best) *best = best_v;
  });
}

// Appends the overlap term S1 (x) S2 (dense n x n each, zero outside the operator's band) as
// the last radial term.
extern "C" int pp_fham_set_overlap(pp_fham* f, const double* S1, const double* S2) {
  return guarded([&] {
    PP_REQUIRE(f && S1 && S2, "null argument");
    FHam& F = f->f;
    PP_REQUIRE(!F.has_overlap, "the operator already has an overlap term");
    const int n = F.n, bw = F.bw, w = 2 * bw + 1;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        PP_REQUIRE(std::abs(i - j) <= bw || (S1[(size_t)i * n + j] == 0.0 && S2[(size_t)i * n + j] == 0.0),
                   "overlap matrix is wider than the operator's band");
    if (F.sep_enabled) {
      // S1 (x) S2 as separable factors after those of the terms: band rows of S1 and S2
      std::vector<double> a1((size_t)n * w, 0.0), b2((size_t)n * w, 0.0);
      for (int i = 0; i < n; ++i)
        for (int jj = 0; jj < w; ++jj) {
          const int j = i - bw + jj;
          if (j < 0 || j >= n) continue;
          a1[(size_t)i * w + jj] = S1[(size_t)i * n + j];
          b2[(size_t)i * w + jj] = S2[(size_t)i * n + j];
        }
      const size_t nsep = (size_t)(F.nterms - F.nband) * n * w;
      auto append = [&](DeviceBuffer<double>& d64, DeviceBuffer<float>& d32, const std::vector<double>& v) {
        std::vector<double> h(nsep + v.size());
        if (nsep) PP_HIP(hipMemcpy(h.data(), d64.ptr, nsep * sizeof(double), hipMemcpyDeviceToHost));
        std::copy(v.begin(), v.end(), h.begin() + nsep);
        d64.upload(h.data(), h.size(), nullptr);
        std::vector<float> hf(h.begin(), h.end());
        d32.upload(hf.data(), hf.size(), nullptr);
        PP_HIP(hipStreamSynchronize(nullptr));
      };
      append(F.sa64, F.sa32, a1);
      append(F.sb64, F.sb32, b2);
      F.has_overlap = true;
      return;
    }
    // dia[i1][jj][d][i2] = S1[i1][i1 - bw + jj] S2[i2][i2 - bw + d]
    std::vector<double> dia((size_t)n * w * w * n, 0.0);
    for (int i1 = 0; i1 < n; ++i1)
      for (int jj = 0; jj < w; ++jj) {
        const int j1 = i1 - bw + jj;
        if (j1 < 0 || j1 >= n) continue;
        const double s1 = S1[(size_t)i1 * n + j1];
        for (int d = 0; d < w; ++d)
          for (int i2 = 0; i2 < n; ++i2) {
            const int j2 = i2 - bw + d;
            if (j2 < 0 || j2 >= n) continue;
            dia[(((size_t)i1 * w + jj) * w + d) * n + i2] = s1 * S2[(size_t)i2 * n + j2];
          }
      }
    const std::vector<double> bt = band_tiles(1, n, bw, dia.data());
    const size_t old_count = F.bt64.count;
    hipStream_t s = nullptr;
    DeviceBuffer<double> b64(old_count + bt.size());
    DeviceBuffer<float> b32(old_count + bt.size());
    PP_HIP(hipMemcpyAsync(b64.ptr, F.bt64.ptr, old_count * sizeof(double), hipMemcpyDeviceToDevice, s));
    PP_HIP(hipMemcpyAsync(b32.ptr, F.bt32.ptr, old_count * sizeof(float), hipMemcpyDeviceToDevice, s));
    PP_HIP(hipMemcpyAsync(b64.ptr + old_count, bt.data(), bt.size() * sizeof(double), hipMemcpyHostToDevice, s));
    std::vector<float> bf(bt.begin(), bt.end());
    PP_HIP(hipMemcpyAsync(b32.ptr + old_count, bf.data(), bf.size() * sizeof(float), hipMemcpyHostToDevice, s));
    PP_HIP(hipStreamSynchronize(s));
    std::swap(F.bt64.ptr, b64.ptr);
    std::swap(F.bt64.count, b64.count);
    std::swap(F.bt32.ptr, b32.ptr);
    std::swap(F.bt32.count, b32.count);
    F.has_overlap = true;
  });
}

extern "C" int pp_fham_multiply_overlap(const pp_fham* f, const double* scalings, const pp_cplx* src, pp_cplx* dst,
                                        const pp_cplx* addend, pp_cplx alpha, pp_cplx beta, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(f && scalings && src && dst, "null argument");
    f->f.apply64(scalings, reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst),
                 reinterpret_cast<const double2*>(addend), make_double2(alpha.re, alpha.im), (hipStream_t)stream,
                 make_double2(beta.re, beta.im));
  });
}

extern "C" int pp_fham_multiply_overlap_f32(const pp_fham* f, const double* scalings, const pp_cplxf* src,
                                            pp_cplxf* dst, const pp_cplxf* addend, pp_cplx alpha, pp_cplx beta,
                                            pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(f && scalings && src && dst, "null argument");
    f->f.apply32(scalings, reinterpret_cast<const float2*>(src), reinterpret_cast<float2*>(dst),
                 reinterpret_cast<const float2*>(addend), make_double2(alpha.re, alpha.im), (hipStream_t)stream,
                 make_double2(beta.re, beta.im));
  });
}

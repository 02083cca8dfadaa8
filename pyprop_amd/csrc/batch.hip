// Batched pAMP propagation of many independent atoms in one persistent launch (no reference
// counterpart for the batching; each atom's step is PamPropagator.AdvanceStep in the
// balanced form, pyprop/propagator/PamPropagator.py:50-59, with pAMP::PropagateVector,
// pamp.h:270-670).
//
// Atom a: psi_a [nl][N] complex128 (reduced spherical l = 0..nl-1 x N B-splines), banded
// overlap S = U^T U (band B), H(t) = H0 + E(t) D with
//   H0: angular-diagonal, radial banded per l   (the Diag x BandNH potential)
//   D : l <-> l-1, l+1 couplings c[l][.] times one banded radial matrix R (Simp x BandNH)
//   E(t) = E0 sin^2(pi t / T) cos(omega t) on [0, T), 0 outside.
// One workgroup owns one atom at a time and runs the whole step:
//   psi <- U psi; Arnoldi (m matvecs of A = U^-T H(t) U^-1, classical Gram-Schmidt, happy
//   breakdown at sqrt(eps)); E = exp(-i dt H_m) by scaling-and-squaring Pade; psi <- |psi| V E[:,0];
//   psi <- U^-1 psi; t += dt.
// The state, the Krylov basis and every reduction are FP64; the banded H multiply runs in
// FP32 (complex64 operands, FP32 accumulation) when F32, else FP64.  Triangular solves keep a
// sliding window of the last B unknowns in registers, one thread per l fibre.
// Workgroups take atoms largest first (order[]), strided by gridDim.x.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <type_traits>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_pade.h"

namespace pp {

struct AtomDesc {
  int nl = 0, N = 0;
  long long psi = 0;  // complex offset of psi [nl][N]
  long long h0 = 0;   // real offset of H0 [nl][w][N] (element (i, i+d-B) at [l][d][i])
  long long r = 0;    // real offset of R [w][N]
  long long cpl = 0;  // real offset of couplings [nl][2] (to l-1, to l+1)
  long long u = 0;    // real offset of U band [N][B+1] (element (i, j), j > i, at [j][B + i - j];
                      // the diagonal slot [i][B] holds 1 / U(i, i))
  long long v = 0;    // complex offset of the Krylov basis [m][nl N]
  double e0 = 0, omega = 0, duration = 0;
};

struct BatchArgs {
  const AtomDesc* atoms;
  const int* order;
  int natoms, m, pade;
  int u_slots;     // doubles reserved for U in LDS (max N (B+1), even)
  double dt, tol;
  double2* psi;
  double2* V;
  double2* err;    // [natom] error estimate E(m-1, 0)
  double* t;       // [natom]
  const double* h0_64;
  const float* h0_32;
  const double* r64;
  const float* r32;
  const double* cpl;
  const double* u;
};

// workgroup sum of a complex value (fixed order; every thread gets the total)
__device__ inline double2 wg_sum2(double2 v, double2* sh) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double2 s = make_double2(0, 0);
  for (int i = 0; i < nw; ++i) {
    s.x += sh[i].x;
    s.y += sh[i].y;
  }
  return s;
}

// Banded triangular solves, one thread per l fibre.  U holds the upper band with the
// diagonal replaced by its reciprocal (element (i, j), j >= i, at [j][B + i - j]).  The
// unknowns of the last B rows sit in registers (win); per row only the term of the newest
// unknown and the final scaling are on the dependency chain: the other B-1 products are
// summed beforehand, and the next row's right-hand side and band are loaded one row ahead.

// x <- U^-1 x (back substitution)
template <int B>
__device__ inline void solve_upper(const double* U, double2* x, int nl, int N) {
  const int l = threadIdx.x;
  if (l < nl) {
    double2* xf = x + (long long)l * N;
    double2 win[B];  // win[d-1] = x[i+d]
#pragma unroll
    for (int d = 0; d < B; ++d) win[d] = make_double2(0, 0);
    double2 rhs = xf[N - 1];
    double u[B + 1];  // u[d] = U(i, i+d) for the current row, u[0] = 1 / U(i, i)
#pragma unroll
    for (int d = 0; d <= B; ++d) u[d] = (N - 1 + d < N) ? U[(N - 1 + d) * (B + 1) + B - d] : 0.0;
    for (int i = N - 1; i >= 0; --i) {
      // prefetch row i-1
      double2 rhs_n = make_double2(0, 0);
      double un[B + 1];
      if (i > 0) {
        rhs_n = xf[i - 1];
#pragma unroll
        for (int d = 0; d <= B; ++d) un[d] = (i - 1 + d < N) ? U[(i - 1 + d) * (B + 1) + B - d] : 0.0;
      }
      // off-chain part: rhs - sum_{d >= 2} u_d x[i+d]
      double2 t = rhs;
#pragma unroll
      for (int d = B; d >= 2; --d) {
        t.x = fma(-u[d], win[d - 1].x, t.x);
        t.y = fma(-u[d], win[d - 1].y, t.y);
      }
      double2 v;
      v.x = fma(-u[1], win[0].x, t.x) * u[0];
      v.y = fma(-u[1], win[0].y, t.y) * u[0];
      xf[i] = v;
#pragma unroll
      for (int d = B - 1; d >= 1; --d) win[d] = win[d - 1];
      win[0] = v;
      rhs = rhs_n;
#pragma unroll
      for (int d = 0; d <= B; ++d) u[d] = un[d];
    }
  }
  __syncthreads();
}

// y <- U^-T y (forward substitution; U real)
template <int B>
__device__ inline void solve_lower_t(const double* U, double2* y, int nl, int N) {
  const int l = threadIdx.x;
  if (l < nl) {
    double2* yf = y + (long long)l * N;
    double2 win[B];  // win[d-1] = y[i-d]
#pragma unroll
    for (int d = 0; d < B; ++d) win[d] = make_double2(0, 0);
    double2 rhs = yf[0];
    double u[B + 1];  // u[d] = U(i-d, i) for the current row, u[0] = 1 / U(i, i)
#pragma unroll
    for (int d = 0; d <= B; ++d) u[d] = (0 - d >= 0) ? U[0 * (B + 1) + B - d] : 0.0;
    for (int i = 0; i < N; ++i) {
      double2 rhs_n = make_double2(0, 0);
      double un[B + 1];
      if (i + 1 < N) {
        rhs_n = yf[i + 1];
#pragma unroll
        for (int d = 0; d <= B; ++d) un[d] = (i + 1 - d >= 0) ? U[(i + 1) * (B + 1) + B - d] : 0.0;
      }
      double2 t = rhs;
#pragma unroll
      for (int d = B; d >= 2; --d) {
        t.x = fma(-u[d], win[d - 1].x, t.x);
        t.y = fma(-u[d], win[d - 1].y, t.y);
      }
      double2 v;
      v.x = fma(-u[1], win[0].x, t.x) * u[0];
      v.y = fma(-u[1], win[0].y, t.y) * u[0];
      yf[i] = v;
#pragma unroll
      for (int d = B - 1; d >= 1; --d) win[d] = win[d - 1];
      win[0] = v;
      rhs = rhs_n;
#pragma unroll
      for (int d = 0; d <= B; ++d) u[d] = un[d];
    }
  }
  __syncthreads();
}

// y = H(t) x (x, y in LDS), FP32 or FP64 arithmetic.  All 2 (2B+1) band loads of an element
// are issued before its FMAs.
template <int B, bool F32>
__device__ inline void ham_apply(const BatchArgs& A, const AtomDesc& d, double E, const double2* x, double2* y) {
  constexpr int w = 2 * B + 1;
  typedef typename std::conditional<F32, float, double>::type R;
  const int nl = d.nl, N = d.N, n = nl * N;
  const double* cpl = A.cpl + d.cpl;
  const R* h0base = F32 ? (const R*)(A.h0_32 + d.h0) : (const R*)(A.h0_64 + d.h0);
  const R* rb = F32 ? (const R*)(A.r32 + d.r) : (const R*)(A.r64 + d.r);
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int l = e / N, i = e - l * N;
    const R* h0 = h0base + (long long)l * w * N;
    R hv[w], rv[w];
#pragma unroll
    for (int dd = 0; dd < w; ++dd) {
      const int j = i + dd - B;
      const bool ok = j >= 0 && j < N;
      hv[dd] = ok ? h0[dd * N + i] : R(0);
      rv[dd] = ok ? rb[dd * N + i] : R(0);
    }
    const R cm = l > 0 ? (R)cpl[2 * l] : R(0), cp = l + 1 < nl ? (R)cpl[2 * l + 1] : R(0);
    R ax = 0, ay = 0, bx = 0, by = 0;
#pragma unroll
    for (int dd = 0; dd < w; ++dd) {
      const int j = i + dd - B;
      if (j < 0 || j >= N) continue;
      const double2 xv = x[(long long)l * N + j];
      ax = fma(hv[dd], (R)xv.x, ax);
      ay = fma(hv[dd], (R)xv.y, ay);
      R sx = 0, sy = 0;
      if (l > 0) {
        const double2 xm = x[(long long)(l - 1) * N + j];
        sx = fma(cm, (R)xm.x, sx);
        sy = fma(cm, (R)xm.y, sy);
      }
      if (l + 1 < nl) {
        const double2 xp = x[(long long)(l + 1) * N + j];
        sx = fma(cp, (R)xp.x, sx);
        sy = fma(cp, (R)xp.y, sy);
      }
      bx = fma(rv[dd], sx, bx);
      by = fma(rv[dd], sy, by);
    }
    const R Er = (R)E;
    y[e] = make_double2((double)fma(Er, bx, ax), (double)fma(Er, by, ay));
  }
  __syncthreads();
}

template <int B, bool F32>
__global__ __launch_bounds__(1024) void batch_pamp_kernel(BatchArgs A) {
  extern __shared__ unsigned char smem[];
  __shared__ double2 red[16];
  __shared__ double2 red8[16][8];
  const int nt = blockDim.x;
  const int m = A.m;
  for (int q = blockIdx.x; q < A.natoms; q += gridDim.x) {
    const int a = A.order[q];
    const AtomDesc d = A.atoms[a];
    const int nl = d.nl, N = d.N, n = nl * N;
    // LDS: U band | Hessenberg m x m | Gram-Schmidt coefficients m | exp column m | pivots m
    //      | X, Y (n each), reused as the 6 m^2 Pade work matrices once the Arnoldi loop ends
    double* U = reinterpret_cast<double*>(smem);
    double2* Hm = reinterpret_cast<double2*>(U + A.u_slots);
    double2* hs = Hm + m * m;
    double2* ex = hs + m;
    int* piv = reinterpret_cast<int*>(ex + m);
    double2* X = ex + 2 * m;
    double2* Y = X + n;
    double2* psi = A.psi + d.psi;
    double2* V = A.V + d.v;
    const double t = A.t[a];
    const double E = (t >= 0 && t < d.duration) ? d.e0 * sin(M_PI * t / d.duration) * sin(M_PI * t / d.duration) *
                                                      cos(d.omega * t)
                                                : 0.0;
    for (int i = threadIdx.x; i < N * (B + 1); i += nt) U[i] = A.u[d.u + i];
    for (int i = threadIdx.x; i < m * m; i += nt) Hm[i] = make_double2(0, 0);
    for (int i = threadIdx.x; i < n; i += nt) Y[i] = psi[i];
    __syncthreads();
    // psi <- U psi (MultiplySqrtOverlap), into X; |psi|
    double2 part = make_double2(0, 0);
    for (int e = threadIdx.x; e < n; e += nt) {
      const int l = e / N, i = e - l * N;
      double2 s = make_double2(0, 0);
      for (int dd = 0; dd <= B && i + dd < N; ++dd) {
        const double uu = dd == 0 ? 1.0 / U[i * (B + 1) + B] : U[(i + dd) * (B + 1) + B - dd];
        const double2 v = Y[(long long)l * N + i + dd];
        s.x = fma(uu, v.x, s.x);
        s.y = fma(uu, v.y, s.y);
      }
      X[e] = s;
      part.x += s.x * s.x + s.y * s.y;
    }
    const double nrm = sqrt(wg_sum2(part, red).x);
    const double inv0 = nrm > 0 ? 1.0 / nrm : 0.0;
    // V0 = psi / |psi| (global) and the operand in X
    for (int e = threadIdx.x; e < n; e += nt) {
      const double2 v = make_double2(X[e].x * inv0, X[e].y * inv0);
      X[e] = v;
      V[e] = v;
    }
    __syncthreads();
    // r = A V0 into Y (X destroyed by the solve)
    solve_upper<B>(U, X, nl, N);
    ham_apply<B, F32>(A, d, E, X, Y);
    solve_lower_t<B>(U, Y, nl, N);
    // H(0,0) = <V0, r>; r -= H(0,0) V0
    {
      double2 p = make_double2(0, 0);
      for (int e = threadIdx.x; e < n; e += nt) {
        const double2 v = V[e], r = Y[e];
        p.x += v.x * r.x + v.y * r.y;
        p.y += v.x * r.y - v.y * r.x;
      }
      const double2 h = wg_sum2(p, red);
      if (threadIdx.x == 0) Hm[0] = h;
      for (int e = threadIdx.x; e < n; e += nt) {
        const double2 v = V[e];
        Y[e].x -= h.x * v.x - h.y * v.y;
        Y[e].y -= h.x * v.y + h.y * v.x;
      }
      __syncthreads();
    }
    int kused = 1;
    for (int j = 0; j < m - 1; ++j) {
      double2 p = make_double2(0, 0);
      for (int e = threadIdx.x; e < n; e += nt) p.x += Y[e].x * Y[e].x + Y[e].y * Y[e].y;
      const double beta = sqrt(wg_sum2(p, red).x);
      if (fabs(beta) < A.tol) break;  // happy breakdown (uniform over the workgroup)
      if (threadIdx.x == 0) Hm[j * m + j + 1] = make_double2(beta, 0);
      const double ib = 1.0 / beta;
      double2* vj1 = V + (long long)(j + 1) * n;
      for (int e = threadIdx.x; e < n; e += nt) {
        const double2 v = make_double2(Y[e].x * ib, Y[e].y * ib);
        X[e] = v;
        vj1[e] = v;
      }
      __syncthreads();
      kused = j + 2;
      solve_upper<B>(U, X, nl, N);
      ham_apply<B, F32>(A, d, E, X, Y);
      solve_lower_t<B>(U, Y, nl, N);
      // classical Gram-Schmidt against V_0..V_{j+1}; H(i, j+1) = h_i.  Eight dot products
      // per pass over the vectors, one workgroup reduction per pass.
      for (int i0 = 0; i0 <= j + 1; i0 += 8) {
        double2 acc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = make_double2(0, 0);
        for (int e = threadIdx.x; e < n; e += nt) {
          const double2 r = Y[e];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            if (i0 + q <= j + 1) {
              const double2 v = V[(long long)(i0 + q) * n + e];
              acc[q].x += v.x * r.x + v.y * r.y;
              acc[q].y += v.x * r.y - v.y * r.x;
            }
          }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
          for (int off = 32; off > 0; off >>= 1) {
            acc[q].x += __shfl_down(acc[q].x, off, 64);
            acc[q].y += __shfl_down(acc[q].y, off, 64);
          }
        const int wv = threadIdx.x >> 6, nw = (nt + 63) >> 6;
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
          for (int q = 0; q < 8; ++q) red8[wv][q] = acc[q];
        }
        __syncthreads();
        if (threadIdx.x < 8 && i0 + (int)threadIdx.x <= j + 1) {
          double2 h = make_double2(0, 0);
          for (int w2 = 0; w2 < nw; ++w2) {
            h.x += red8[w2][threadIdx.x].x;
            h.y += red8[w2][threadIdx.x].y;
          }
          Hm[(j + 1) * m + i0 + threadIdx.x] = h;
          hs[i0 + threadIdx.x] = h;
        }
        __syncthreads();
      }
      __syncthreads();
      {
        for (int e = threadIdx.x; e < n; e += nt) {
          double2 r = Y[e];
          for (int i = 0; i <= j + 1; ++i) {
            const double2 v = V[(long long)i * n + e], h = hs[i];
            r.x -= h.x * v.x - h.y * v.y;
            r.y -= h.x * v.y + h.y * v.x;
          }
          Y[e] = r;
        }
        __syncthreads();
      }
    }
    // E = exp(-i dt H_m) (work matrices over X, Y); psi = |psi| V E[:, 0]; psi <- U^-1 psi
    __syncthreads();
    pade_exp_wg(Hm, m, make_double2(A.dt, 0.0), A.pade, X, ex, A.err + a, piv);
    for (int e = threadIdx.x; e < n; e += nt) {
      double2 s = make_double2(0, 0);
      for (int i = 0; i < kused; ++i) {
        const double2 v = V[(long long)i * n + e], c = ex[i];
        s.x += c.x * v.x - c.y * v.y;
        s.y += c.x * v.y + c.y * v.x;
      }
      X[e] = make_double2(s.x * nrm, s.y * nrm);
    }
    __syncthreads();
    solve_upper<B>(U, X, nl, N);
    for (int e = threadIdx.x; e < n; e += nt) psi[e] = X[e];
    if (threadIdx.x == 0) A.t[a] = t + A.dt;
    __syncthreads();
  }
}

}  // namespace pp

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
namespace pp {

void band_cholesky_upper_host(int n, int sd, const double* s, std::vector<double>& u);  // overlap.hip

struct Batch {
  int natoms = 0, band = 0, m = 20, pade = 6;
  bool f32 = true;
  int threads = 256;  // workgroup size (autotuned)
  int grid = 0;       // persistent grid (<= natoms)
  int u_slots = 0;
  size_t lds = 0;
  std::vector<AtomDesc> host_atoms;
  DeviceBuffer<AtomDesc> atoms;
  DeviceBuffer<int> order;
  DeviceBuffer<double2> psi, V, err;
  DeviceBuffer<double> t, h0_64, r64, cpl, u;
  DeviceBuffer<float> h0_32, r32;
  long long psi_total = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_steps = 0;
  std::vector<double> tune_ms;
  ~Batch() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  BatchArgs args(double dt) const {
    BatchArgs a;
    a.atoms = atoms.ptr;
    a.order = order.ptr;
    a.natoms = natoms;
    a.m = m;
    a.pade = pade;
    a.u_slots = u_slots;
    a.dt = dt;
    a.tol = std::sqrt(std::numeric_limits<double>::epsilon());
    a.psi = psi.ptr;
    a.V = V.ptr;
    a.err = err.ptr;
    a.t = t.ptr;
    a.h0_64 = h0_64.ptr;
    a.h0_32 = h0_32.ptr;
    a.r64 = r64.ptr;
    a.r32 = r32.ptr;
    a.cpl = cpl.ptr;
    a.u = u.ptr;
    return a;
  }
  void launch(double dt, int thr, hipStream_t s) const;
};

template <int B>
static void launch_b(const Batch& b, double dt, int thr, hipStream_t s) {
  const BatchArgs a = b.args(dt);
  if (b.f32)
    hipLaunchKernelGGL((batch_pamp_kernel<B, true>), dim3(b.grid), dim3(thr), b.lds, s, a);
  else
    hipLaunchKernelGGL((batch_pamp_kernel<B, false>), dim3(b.grid), dim3(thr), b.lds, s, a);
}

template <int B>
static void set_lds_b(size_t lds) {
  PP_HIP(hipFuncSetAttribute((const void*)batch_pamp_kernel<B, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds));
  PP_HIP(hipFuncSetAttribute((const void*)batch_pamp_kernel<B, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds));
}

void Batch::launch(double dt, int thr, hipStream_t s) const {
  switch (band) {
    case 4: launch_b<4>(*this, dt, thr, s); break;
    case 5: launch_b<5>(*this, dt, thr, s); break;
    case 6: launch_b<6>(*this, dt, thr, s); break;
    case 7: launch_b<7>(*this, dt, thr, s); break;
    case 8: launch_b<8>(*this, dt, thr, s); break;
    default: throw Error(PP_ERR_INVALID, "batched pAMP supports B-spline orders 5..9");
  }
  PP_HIP(hipGetLastError());
}

}  // namespace pp

struct pp_batch {
  pp::Batch b;
};

using namespace pp;

extern "C" int pp_batch_create(int natoms, const int* nl, const int* N, int band, const double* h0,
                               const double* r, const double* cpl, const double* s_upper, const double* pulse,
                               int krylov, int pade_order, int fp32_matvec, pp_batch** out) {
  return guarded([&] {
    PP_REQUIRE(out && nl && N && h0 && r && cpl && s_upper && pulse, "null argument");
    PP_REQUIRE(natoms > 0 && band >= 4 && band <= 8, "batched pAMP supports band half-widths 4..8");
    PP_REQUIRE(krylov >= 2 && krylov <= 64 && pade_order >= 1 && pade_order < 30, "invalid Krylov/Pade size");
    auto* p = new pp_batch();
    Batch& b = p->b;
    b.natoms = natoms;
    b.band = band;
    b.m = krylov;
    b.pade = pade_order;
    b.f32 = fp32_matvec != 0;
    const int w = 2 * band + 1;
    long long o_psi = 0, o_h0 = 0, o_r = 0, o_c = 0, o_u = 0, o_v = 0;
    int max_n = 0, max_N = 0;
    std::vector<double> ubuf;
    for (int a = 0; a < natoms; ++a) {
      PP_REQUIRE(nl[a] >= 1 && nl[a] <= 64 && N[a] > 2 * band && N[a] <= 512, "atom size out of range");
      AtomDesc d;
      d.nl = nl[a];
      d.N = N[a];
      d.psi = o_psi;
      d.h0 = o_h0;
      d.r = o_r;
      d.cpl = o_c;
      d.u = o_u;
      d.v = o_v;
      d.e0 = pulse[3 * a];
      d.omega = pulse[3 * a + 1];
      d.duration = pulse[3 * a + 2];
      const long long n = (long long)d.nl * d.N;
      PP_REQUIRE(d.nl <= 1024, "too many channels");
      o_psi += n;
      o_h0 += (long long)d.nl * w * d.N;
      o_r += (long long)w * d.N;
      o_c += 2LL * d.nl;
      o_u += (long long)d.N * (band + 1);
      o_v += (long long)krylov * n;
      max_n = std::max<int>(max_n, (int)n);
      max_N = std::max(max_N, d.N);
      b.host_atoms.push_back(d);
    }
    b.psi_total = o_psi;
    // Cholesky factor of every atom's overlap
    ubuf.resize(o_u);
    for (int a = 0; a < natoms; ++a) {
      const AtomDesc& d = b.host_atoms[a];
      std::vector<double> uu;
      band_cholesky_upper_host(d.N, band, s_upper + d.u, uu);
      for (int i = 0; i < d.N; ++i) uu[(size_t)i * (band + 1) + band] = 1.0 / uu[(size_t)i * (band + 1) + band];
      std::copy(uu.begin(), uu.end(), ubuf.begin() + d.u);
    }
    b.u_slots = (max_N * (band + 1) + 1) & ~1;
    b.lds = (size_t)b.u_slots * sizeof(double) + ((size_t)krylov * krylov + 4 * krylov) * sizeof(double2) +
            (size_t)std::max<long long>(2LL * max_n, 6LL * krylov * krylov) * sizeof(double2);
    PP_REQUIRE(b.lds <= 140 * 1024, "largest atom does not fit the workgroup's LDS (about nl N <= 3900)");
    // largest atoms first
    std::vector<int> ord(natoms);
    for (int a = 0; a < natoms; ++a) ord[a] = a;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
      return (long long)b.host_atoms[x].nl * b.host_atoms[x].N > (long long)b.host_atoms[y].nl * b.host_atoms[y].N;
    });
    hipStream_t s = nullptr;
    try {
      b.atoms.upload(b.host_atoms.data(), natoms, s);
      b.order.upload(ord.data(), natoms, s);
      b.h0_64.upload(h0, o_h0, s);
      std::vector<float> f(h0, h0 + o_h0);
      b.h0_32.upload(f.data(), o_h0, s);
      b.r64.upload(r, o_r, s);
      std::vector<float> fr(r, r + o_r);
      b.r32.upload(fr.data(), o_r, s);
      b.cpl.upload(cpl, o_c, s);
      b.u.upload(ubuf.data(), o_u, s);
      b.psi.resize(o_psi);
      b.V.resize(o_v);
      b.err.resize(natoms);
      std::vector<double> zt(natoms, 0.0);
      b.t.upload(zt.data(), natoms, s);
      switch (band) {
        case 4: set_lds_b<4>(b.lds); break;
        case 5: set_lds_b<5>(b.lds); break;
        case 6: set_lds_b<6>(b.lds); break;
        case 7: set_lds_b<7>(b.lds); break;
        case 8: set_lds_b<8>(b.lds); break;
      }
      PP_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete p;
      throw;
    }
    int ncu = 0;
    int dev = 0;
    PP_HIP(hipGetDevice(&dev));
    PP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    b.grid = std::min(natoms, std::max(1, ncu));
    *out = p;
  });
}

extern "C" int pp_batch_destroy(pp_batch* b) {
  return guarded([&] { delete b; });
}

extern "C" int pp_batch_set_state(pp_batch* p, const pp_cplx* psi, const double* t) {
  return guarded([&] {
    PP_REQUIRE(p && psi && t, "null argument");
    Batch& b = p->b;
    // a step or graph replay may still run on a non-blocking stream: wait for it before
    // overwriting the state it reads and writes
    PP_HIP(hipDeviceSynchronize());
    PP_HIP(hipMemcpy(b.psi.ptr, psi, b.psi_total * sizeof(double2), hipMemcpyHostToDevice));
    PP_HIP(hipMemcpy(b.t.ptr, t, b.natoms * sizeof(double), hipMemcpyHostToDevice));
  });
}

extern "C" int pp_batch_get_state(const pp_batch* p, pp_cplx* psi, double* t, pp_cplx* err) {
  return guarded([&] {
    PP_REQUIRE(p, "null argument");
    const Batch& b = p->b;
    PP_HIP(hipDeviceSynchronize());
    if (psi) PP_HIP(hipMemcpy(psi, b.psi.ptr, b.psi_total * sizeof(double2), hipMemcpyDeviceToHost));
    if (t) PP_HIP(hipMemcpy(t, b.t.ptr, b.natoms * sizeof(double), hipMemcpyDeviceToHost));
    if (err) PP_HIP(hipMemcpy(err, b.err.ptr, b.natoms * sizeof(double2), hipMemcpyDeviceToHost));
  });
}

// Autotuner: time each workgroup size on `trial_steps` steps from the current state, restore
// the state, keep the fastest.  timings_ms (host, 5 entries, may be NULL) receives the
// per-step time of 64, 128, 256, 512, 1024 threads.
extern "C" int pp_batch_autotune(pp_batch* p, double dt, int trial_steps, int* best_threads, double* timings_ms,
                                 pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && trial_steps >= 1, "invalid argument");
    Batch& b = p->b;
    hipStream_t s = (hipStream_t)stream;
    DeviceBuffer<double2> psi_save;
    DeviceBuffer<double> t_save;
    psi_save.resize(b.psi_total);
    t_save.resize(b.natoms);
    PP_HIP(hipMemcpyAsync(psi_save.ptr, b.psi.ptr, b.psi_total * sizeof(double2), hipMemcpyDeviceToDevice, s));
    PP_HIP(hipMemcpyAsync(t_save.ptr, b.t.ptr, b.natoms * sizeof(double), hipMemcpyDeviceToDevice, s));
    const int cands[5] = {64, 128, 256, 512, 1024};
    hipEvent_t e0, e1;
    PP_HIP(hipEventCreate(&e0));
    PP_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int best_thr = b.threads;
    b.tune_ms.assign(5, 0.0);
    for (int c = 0; c < 5; ++c) {
      const int thr = cands[c];
      b.launch(dt, thr, s);  // warm-up
      PP_HIP(hipEventRecord(e0, s));
      for (int k = 0; k < trial_steps; ++k) b.launch(dt, thr, s);
      PP_HIP(hipEventRecord(e1, s));
      PP_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      PP_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double per = ms / trial_steps;
      b.tune_ms[c] = per;
      if (timings_ms) timings_ms[c] = per;
      if (per < best) {
        best = per;
        best_thr = thr;
      }
      PP_HIP(hipMemcpyAsync(b.psi.ptr, psi_save.ptr, b.psi_total * sizeof(double2), hipMemcpyDeviceToDevice, s));
      PP_HIP(hipMemcpyAsync(b.t.ptr, t_save.ptr, b.natoms * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    PP_HIP(hipStreamSynchronize(s));
    PP_HIP(hipEventDestroy(e0));
    PP_HIP(hipEventDestroy(e1));
    b.threads = best_thr;
    if (best_threads) *best_threads = best_thr;
  });
}

extern "C" int pp_batch_set_threads(pp_batch* p, int threads) {
  return guarded([&] {
    PP_REQUIRE(p && (threads == 64 || threads == 128 || threads == 256 || threads == 512 || threads == 1024),
               "threads must be 64, 128, 256, 512 or 1024");
    p->b.threads = threads;
  });
}

// nsteps launches of the persistent step kernel on `stream`
extern "C" int pp_batch_step(pp_batch* p, double dt, int nsteps, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && nsteps >= 0, "invalid argument");
    for (int k = 0; k < nsteps; ++k) p->b.launch(dt, p->b.threads, (hipStream_t)stream);
  });
}

// capture `steps_per_graph` step launches into a hipGraph (replaced on every call)
extern "C" int pp_batch_capture(pp_batch* p, double dt, int steps_per_graph, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && steps_per_graph >= 1, "invalid argument");
    Batch& b = p->b;
    hipStream_t s = (hipStream_t)stream;
    PP_REQUIRE(s != nullptr, "graph capture needs a non-default stream");
    if (b.graph_exec) {
      PP_HIP(hipGraphExecDestroy(b.graph_exec));
      b.graph_exec = nullptr;
    }
    if (b.graph) {
      PP_HIP(hipGraphDestroy(b.graph));
      b.graph = nullptr;
    }
    PP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < steps_per_graph; ++k) b.launch(dt, b.threads, s);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    PP_HIP(hipStreamEndCapture(s, &b.graph));
    PP_HIP(hipGraphInstantiate(&b.graph_exec, b.graph, nullptr, nullptr, 0));
    b.graph_steps = steps_per_graph;
  });
}

extern "C" int pp_batch_replay(pp_batch* p, int replays, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && p->b.graph_exec && replays >= 0, "no captured graph");
    for (int k = 0; k < replays; ++k) PP_HIP(hipGraphLaunch(p->b.graph_exec, (hipStream_t)stream));
  });
}

extern "C" int pp_batch_info(const pp_batch* p, int* threads, int* grid, long long* psi_elements,
                             long long* lds_bytes) {
  return guarded([&] {
    PP_REQUIRE(p, "null argument");
    if (threads) *threads = p->b.threads;
    if (grid) *grid = p->b.grid;
    if (psi_elements) *psi_elements = p->b.psi_total;
    if (lds_bytes) *lds_bytes = (long long)p->b.lds;
  });
}

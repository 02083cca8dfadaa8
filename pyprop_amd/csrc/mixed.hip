// Mixed-precision Cayley step: complex64 inner solves (FP32 vectors and arithmetic, FP64
// accumulation of inner products), FP64 outer iterative refinement.
//
// No reference counterpart; it solves the system of
// pyprop/propagator/CayleyPropagator.py:71-150,
//   (S + i dt/2 H(t+dt/2)) psi(t+dt) = (S - i dt/2 H(t+dt/2)) psi(t),
// by
//   b = S psi - i a H psi,  x = psi,  r = b - A x = -2 i a H psi        (a = dt/2, one FP64 multiply)
//   repeat:  stop if |r| <= tol_rel |b|
//            d ~= A^-1 r by flexible GMRES in FP32, right preconditioner P_j
//            x += d;  r = b - A x                                       (FP64)
// Flexible GMRES stores the preconditioned directions z_j, so P_j may change from one
// iteration to the next and may be nonlinear (the learned correction is used on the first
// direction only).
//
// Kronecker preconditioner: the exact inverse of the separable part of each channel block,
//   K_c = S x S + a (h_l1 x S + S x h_l2),  h_l E_l = S E_l diag(lam_l),  E_l^T S E_l = I
//   K_c^-1 B = E_l1 [ (E_l1^T B E_l2) / (1 + a (lam_l1,i + lam_l2,j)) ] E_l2^T,
// four n x n products per channel, one workgroup per channel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"
#include "pp_mixed.h"
#include "pp_learned.h"
#include "pp_factored.h"

namespace pp {

static int grid_n(long long n) {
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

// ------------------------------------------------------------------------------------
// FP32 vector kernels (FP64 accumulation in reductions)
// ------------------------------------------------------------------------------------
// y = x / sqrt(*n2) in FP32 (0 where *n2 is 0): the squared norm is read on the device,
// from the finishing kernel of the reduction that formed it
__global__ void f64_to_f32_unit_kernel(long long n, const double2* __restrict__ n2, const double2* __restrict__ x,
                                       float2* __restrict__ y) {
  const double q = n2->x;
  const double sc = q > 0 ? 1.0 / sqrt(q) : 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = make_float2((float)(sc * x[i].x), (float)(sc * x[i].y));
}



// Sum over the blocks of partial[b][j] (j < count, partial rows of `width`) in a fixed order:
// 8 threads per column stride over the blocks, each issuing the loads of U blocks at once and
// adding them in block order into one accumulator, then a shuffle tree.  Every caller uses
// this function with 256 or more threads per block, so any block that evaluates it, and
// finish_partials_kernel, obtains the same bits.  dst may be LDS or global memory; the caller
// synchronises.  Loads past the last block read the last block's row and are multiplied by 0
// (no load behind a branch).
__device__ __forceinline__ void sum_partials(const double2* __restrict__ partial, int blocks, int width, int count,
                                             double2* dst) {
  constexpr int U = 12;
  const int q = threadIdx.x & 7;
  for (int j0 = 0; j0 < count; j0 += 32) {
    const int j = j0 + (int)(threadIdx.x >> 3);
    double2 acc = make_double2(0, 0);
    if (j < count) {
      for (int b0 = q; b0 < blocks; b0 += 8 * U) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int b = b0 + 8 * u;
          v[u] = partial[(long long)min(b, blocks - 1) * width + j];
        }
        __builtin_amdgcn_sched_barrier(0);  // all loads issued before the first add
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double k = b0 + 8 * u < blocks ? 1.0 : 0.0;
          acc.x += k * v[u].x;
          acc.y += k * v[u].y;
        }
      }
    }
    for (int off = 4; off > 0; off >>= 1) {
      acc.x += __shfl_down(acc.x, off, 8);
      acc.y += __shfl_down(acc.y, off, 8);
    }
    if (q == 0 && j < count) dst[j] = acc;
  }
}

// out = [sum of p1 (c1 columns), sum of p2 (c2 columns), sum of p3 (c3 columns)], one block;
// out2 (device memory, may be NULL) receives the same values, for kernels that follow
__global__ __launch_bounds__(256) void finish_partials_kernel(int blocks, const double2* __restrict__ p1, int w1,
                                                              int c1, const double2* __restrict__ p2, int w2, int c2,
                                                              const double2* __restrict__ p3, int w3, int c3,
                                                              double2* __restrict__ out, double2* __restrict__ out2) {
  __shared__ double2 res[128];
  if (c1) sum_partials(p1, blocks, w1, c1, res);
  if (c2) sum_partials(p2, blocks, w2, c2, res + c1);
  if (c3) sum_partials(p3, blocks, w3, c3, res + c1 + c2);
  __syncthreads();
  for (int j = threadIdx.x; j < c1 + c2 + c3; j += blockDim.x) {
    out[j] = res[j];
    if (out2) out2[j] = res[j];
  }
}

// One pass of classical Gram-Schmidt over FP32 vectors with FP64 sums; block partials go to
// partial[block][*] (finished by the next kernel or finish_partials_kernel):
//   UPDATE: h = the finished previous dots (every block sums hpart itself, in the same order),
//           then w -= sum_{j<m} V_j h_j (applied in FP32)
//   DOT:    partial[b][j] = V_j^H w (j < m), partial[b][m] = |w|^2, after the update;
//           else partial[b][0] = |w|^2
// The update and the dots of a block cover the same slice with the same thread-to-element
// map, so each thread reads back only what it wrote.  wout (UPDATE without DOT, may be NULL):
// the updated vector goes there instead of back into w.  A thread takes its elements E at a
// time and issues all their loads (8 basis vectors per batch) before using any: the indices
// are clamped into the slice and the basis rows past m carry the coefficient 0, so no load
// sits behind a branch.
template <bool UPDATE, bool DOT>
__global__ __launch_bounds__(256) void cgs32_kernel(long long n, int m, const float2* __restrict__ V, long long ld,
                                                    const double2* __restrict__ hpart, int hwidth,
                                                    float2* __restrict__ w, double2* __restrict__ partial,
                                                    float2* __restrict__ wout = nullptr) {
  constexpr int E = 4;  // elements per thread per pass
  __shared__ double2 h64s[128];
  __shared__ float2 hs[136];
  __shared__ double2 sh[4][8];
  const int width = DOT ? m + 1 : 1;
  const long long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long long beg = chunk * blockIdx.x;
  const long long end = min(n, beg + chunk);
  double nrm = 0;
  if (UPDATE) {
    sum_partials(hpart, gridDim.x, hwidth, m, h64s);
    __syncthreads();
    for (int j = threadIdx.x; j < 136; j += blockDim.x)
      hs[j] = j < m ? make_float2((float)h64s[j].x, (float)h64s[j].y) : make_float2(0.f, 0.f);
    __syncthreads();
    for (long long i0 = beg + threadIdx.x; i0 < end; i0 += E * 256) {
      long long ie[E];
#pragma unroll
      for (int e = 0; e < E; ++e) ie[e] = min(i0 + e * 256, end - 1);
      float2 t[E];
#pragma unroll
      for (int e = 0; e < E; ++e) t[e] = make_float2(0.f, 0.f);
      for (int j0 = 0; j0 < m; j0 += 8) {
        float2 v[8][E];
#pragma unroll
        for (int g = 0; g < 8; ++g)
#pragma unroll
          for (int e = 0; e < E; ++e) v[g][e] = V[(long long)min(j0 + g, m - 1) * ld + ie[e]];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int g = 0; g < 8; ++g) {
          const float2 h = hs[j0 + g];  // 0 past m
#pragma unroll
          for (int e = 0; e < E; ++e) {
            t[e].x = fmaf(v[g][e].x, h.x, t[e].x);
            t[e].x = fmaf(-v[g][e].y, h.y, t[e].x);
            t[e].y = fmaf(v[g][e].x, h.y, t[e].y);
            t[e].y = fmaf(v[g][e].y, h.x, t[e].y);
          }
        }
      }
      float2 wv[E];
#pragma unroll
      for (int e = 0; e < E; ++e) wv[e] = w[ie[e]];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const bool ok = i0 + e * 256 < end;
        wv[e].x -= t[e].x;
        wv[e].y -= t[e].y;
        nrm += ok ? (double)wv[e].x * wv[e].x + (double)wv[e].y * wv[e].y : 0.0;
      }
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (i0 + e * 256 >= end) continue;
        if (!DOT && wout) wout[ie[e]] = wv[e];
        else w[ie[e]] = wv[e];
      }
    }
  }
  if (DOT) {
    for (int j0 = 0; j0 <= m; j0 += 8) {
      double2 acc[8];
#pragma unroll
      for (int g = 0; g < 8; ++g) acc[g] = make_double2(0, 0);
      for (long long i0 = beg + threadIdx.x; i0 < end; i0 += E * 256) {
        long long ie[E];
#pragma unroll
        for (int e = 0; e < E; ++e) ie[e] = min(i0 + e * 256, end - 1);
        float2 wv[E], v[8][E];
#pragma unroll
        for (int e = 0; e < E; ++e) wv[e] = w[ie[e]];
#pragma unroll
        for (int g = 0; g < 8; ++g)
#pragma unroll
          for (int e = 0; e < E; ++e) v[g][e] = V[(long long)min(j0 + g, max(m - 1, 0)) * ld + ie[e]];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < E; ++e) {  // selects, not branches (a load used in a branch is sunk into it)
          const bool ok = i0 + e * 256 < end;
          const double nn = (double)wv[e].x * wv[e].x + (double)wv[e].y * wv[e].y;
#pragma unroll
          for (int g = 0; g < 8; ++g) {
            const int j = j0 + g;
            const double dx = (double)v[g][e].x * wv[e].x + (double)v[g][e].y * wv[e].y;
            const double dy = (double)v[g][e].x * wv[e].y - (double)v[g][e].y * wv[e].x;
            acc[g].x += ok ? (j < m ? dx : j == m ? nn : 0.0) : 0.0;
            acc[g].y += ok && j < m ? dy : 0.0;
          }
        }
      }
      // block sums of the batch: shuffle trees per wave, one barrier, then the four waves'
      // sums in wave order
#pragma unroll
      for (int g = 0; g < 8; ++g)
        for (int off = 32; off > 0; off >>= 1) {
          acc[g].x += __shfl_down(acc[g].x, off, 64);
          acc[g].y += __shfl_down(acc[g].y, off, 64);
        }
      __syncthreads();  // the previous batch's sh is read
      if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int g = 0; g < 8; ++g) sh[threadIdx.x >> 6][g] = acc[g];
      }
      __syncthreads();
      if (threadIdx.x < 8 && j0 + (int)threadIdx.x <= m) {
        const int g = threadIdx.x;
        double2 t = sh[0][g];
        for (int qq = 1; qq < 4; ++qq) {
          t.x += sh[qq][g].x;
          t.y += sh[qq][g].y;
        }
        partial[(long long)blockIdx.x * width + j0 + g] = t;
      }
    }
  } else {
    if (!UPDATE)
      for (long long i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const float2 wv = w[i];
        nrm += (double)wv.x * wv.x + (double)wv.y * wv.y;
      }
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][0] = make_double2(nrm, 0.0);
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(sh[0][0].x + sh[1][0].x + sh[2][0].x + sh[3][0].x, 0.0);
  }
}

// One-pass Gram-Schmidt finish: out = [h_0 .. h_{m-1}, |w|^2 before the update | m + 1
// zeros | |w|^2 after the update] (the layout of the two-pass results, the second dots zero)
__global__ __launch_bounds__(256) void finish_cgs1_kernel(int blocks, const double2* __restrict__ p1, int m,
                                                          const double2* __restrict__ p3, double2* __restrict__ out,
                                                          double2* __restrict__ out2) {
  __shared__ double2 res[130];
  sum_partials(p1, blocks, m + 1, m + 1, res);
  sum_partials(p3, blocks, 1, 1, res + m + 1);
  __syncthreads();
  const int width = m + 1;
  for (int j = threadIdx.x; j < 2 * width + 1; j += blockDim.x) {
    const double2 v = j < width ? res[j] : (j == 2 * width ? res[width] : make_double2(0.0, 0.0));
    out[j] = v;
    if (out2) out2[j] = v;
  }
}

// r = b - r (r holds A x on entry); partial[block] = block's |r|^2
__global__ __launch_bounds__(256) void residual_partial_kernel(long long n, const double2* __restrict__ b,
                                                               double2* __restrict__ r,
                                                               double2* __restrict__ partial) {
  __shared__ double sh[4];
  double acc = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double2 v = make_double2(b[i].x - r[i].x, b[i].y - r[i].y);
    r[i] = v;
    acc = fma(v.x, v.x, acc);
    acc = fma(v.y, v.y, acc);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(sh[0] + sh[1] + sh[2] + sh[3], 0.0);
}

// out[j] = (sum_b part[2 b + j], 0) for j < count (<= 2), one block: every thread sums its
// blocks (all loads issued first) in order, then a shuffle tree and the four waves in order;
// out2 (device memory, may be NULL) receives the same values
__global__ __launch_bounds__(256) void finish_norms_kernel(int blocks, const double2* __restrict__ part, int count,
                                                           double2* __restrict__ out, double2* __restrict__ out2) {
  constexpr int U = 8;
  __shared__ double sh[2][4];
  double a0 = 0, a1 = 0;
  for (int b0 = threadIdx.x; b0 < blocks; b0 += 256 * U) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = part[min(b0 + 256 * u, blocks - 1)];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double k = b0 + 256 * u < blocks ? 1.0 : 0.0;
      a0 += k * v[u].x;
      a1 += k * v[u].y;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_down(a0, off, 64);
    a1 += __shfl_down(a1, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a0;
    sh[1][threadIdx.x >> 6] = a1;
  }
  __syncthreads();
  if (threadIdx.x < count) {
    const int j = threadIdx.x;
    const double2 r = make_double2(((sh[j][0] + sh[j][1]) + sh[j][2]) + sh[j][3], 0.0);
    out[j] = r;
    if (out2) out2[j] = r;
  }
}

// blocks of the fused reductions: about 1024 elements each
static int fused_blocks(long long n) {
  long long b = (n + 1023) / 1024;
  if (b < 1) b = 1;
  if (b > 512) b = 512;
  return (int)b;
}

// y = x / sqrt(n2) with the squared norm on the device (a zero vector stays zero)
__global__ void unit32_to_kernel(long long n, const double2* __restrict__ n2, const float2* __restrict__ x,
                                 float2* __restrict__ y) {
  const double q = n2->x;
  const float sc = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = make_float2(sc * x[i].x, sc * x[i].y);
}

__global__ void scale32_kernel(long long n, float s, const float2* __restrict__ x, float2* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = make_float2(s * x[i].x, s * x[i].y);
}

// out = sum_j c_j Z_j (FP32 vectors; the coefficients travel in the kernel arguments)
struct LincombCoef {
  float2 c[120];
};

constexpr int kDevLsMax = 16;  // inner iterations of a correction solved on the device

// x = base + sum_j c_j Z_j (the combination in FP32, added in FP64; base may alias x)
__global__ void lincomb_add32_kernel(long long n, int m, const float2* __restrict__ Z, long long ld, LincombCoef cs,
                                     const double2* base, double2* x) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float2 t = make_float2(0.f, 0.f);
    for (int j = 0; j < m; ++j) {
      const float2 v = Z[j * ld + i];
      t.x = fmaf(v.x, cs.c[j].x, t.x);
      t.x = fmaf(-v.y, cs.c[j].y, t.x);
      t.y = fmaf(v.x, cs.c[j].y, t.y);
      t.y = fmaf(v.y, cs.c[j].x, t.y);
    }
    const double2 b = base[i];
    x[i] = make_double2(b.x + (double)t.x, b.y + (double)t.y);
  }
}

// The same combination with the coefficients on the device (m <= kDevLsMax).
__global__ void lincomb_add32_dev_kernel(long long n, int m, const float2* __restrict__ Z, long long ld,
                                         const float2* __restrict__ cs, const double2* base, double2* x) {
  __shared__ float2 c[kDevLsMax];
  if (threadIdx.x < kDevLsMax) c[threadIdx.x] = threadIdx.x < m ? cs[threadIdx.x] : make_float2(0.f, 0.f);
  __syncthreads();
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float2 t = make_float2(0.f, 0.f);
    for (int j = 0; j < m; ++j) {
      const float2 v = Z[j * ld + i];
      t.x = fmaf(v.x, c[j].x, t.x);
      t.x = fmaf(-v.y, c[j].y, t.x);
      t.y = fmaf(v.x, c[j].y, t.y);
      t.y = fmaf(v.y, c[j].x, t.y);
    }
    const double2 b = base[i];
    x[i] = make_double2(b.x + (double)t.x, b.y + (double)t.y);
  }
}

__device__ inline double2 zmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 zscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }

// The inner least-squares problem min |beta e1 - Hbar y| on the device, from the
// Gram-Schmidt results of k iterations (hres + j * stride holds iteration j's first dots and
// |w|^2, second dots and |w|^2, final |w|^2: Hbar(i, j) = first + second dots, Hbar(j+1, j) =
// |w| after both passes, the sums the host forms) and beta^2 = beta2->x.  Givens QR in
// FP64 on one thread.  y -> ycoef (FP32, for the combination); the residual estimate
// |g_{j+1}| / beta after each column j -> rel[j].x.  A column that the rotations leave
// (numerically) zero, as after a breakdown, gets coefficient 0.
__global__ void hessenberg_lsq_kernel(int k, int stride, const double2* __restrict__ hres,
                                      const double2* __restrict__ beta2, float2* __restrict__ ycoef,
                                      double2* __restrict__ rel) {
  // the work arrays live in LDS (one thread uses them), not in per-lane scratch memory
  __shared__ double2 Rm[kDevLsMax][kDevLsMax];
  __shared__ double2 g[kDevLsMax + 1];
  __shared__ double cs[kDevLsMax];
  __shared__ double2 sn[kDevLsMax];
  __shared__ double2 col[kDevLsMax + 1];
  __shared__ double2 y[kDevLsMax];
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double beta = sqrt(fmax(0.0, beta2->x));
  for (int i = 0; i <= k; ++i) g[i] = make_double2(i == 0 ? beta : 0.0, 0.0);
  double scale = 0.0;
  for (int j = 0; j < k; ++j) {
    const double2* hs = hres + (long long)j * stride;
    const int w1 = j + 2;
    for (int i = 0; i <= j; ++i) col[i] = zadd(hs[i], hs[w1 + i]);
    col[j + 1] = make_double2(sqrt(fmax(0.0, hs[2 * w1].x)), 0.0);
    for (int i = 0; i <= j + 1; ++i) scale = fmax(scale, fabs(col[i].x) + fabs(col[i].y));
    for (int i = 0; i < j; ++i) {
      const double2 x = col[i], y = col[i + 1];
      col[i] = zadd(zscale(cs[i], x), zmul(sn[i], y));
      col[i + 1] = zadd(zscale(-1.0, zmul(zconj(sn[i]), x)), zscale(cs[i], y));
    }
    const double2 a = col[j], b = col[j + 1];
    const double na = hypot(a.x, a.y), nb = hypot(b.x, b.y), rho = hypot(na, nb);
    if (rho == 0.0) {
      cs[j] = 1.0;
      sn[j] = make_double2(0.0, 0.0);
    } else if (na == 0.0) {
      cs[j] = 0.0;
      sn[j] = zscale(1.0 / nb, zconj(b));
    } else {
      cs[j] = na / rho;
      sn[j] = zscale(1.0 / (na * rho), zmul(a, zconj(b)));
    }
    col[j] = zadd(zscale(cs[j], a), zmul(sn[j], b));
    for (int i = 0; i <= j; ++i) Rm[i][j] = col[i];
    const double2 x = g[j];
    g[j] = zscale(cs[j], x);
    g[j + 1] = zscale(-1.0, zmul(zconj(sn[j]), x));
    rel[j] = make_double2(beta > 0 ? hypot(g[j + 1].x, g[j + 1].y) / beta : 0.0, 0.0);
  }
  for (int i = k - 1; i >= 0; --i) {
    double2 t = g[i];
    for (int l = i + 1; l < k; ++l) t = zadd(t, zscale(-1.0, zmul(Rm[i][l], y[l])));
    const double2 d = Rm[i][i];
    const double dd = d.x * d.x + d.y * d.y;
    y[i] = (dd > 0.0 && sqrt(dd) > 1e-14 * scale) ? zscale(1.0 / dd, zmul(t, zconj(d))) : make_double2(0.0, 0.0);
  }
  for (int i = 0; i < k; ++i) ycoef[i] = make_float2((float)y[i].x, (float)y[i].y);
}

// partial[block][0] = |a|^2, partial[block][1] = |b|^2 over the block's share (FP64)
__global__ __launch_bounds__(256) void norm2_pair_partial_kernel(long long n, const double2* __restrict__ a,
                                                                 const double2* __restrict__ b,
                                                                 double2* __restrict__ partial) {
  __shared__ double sh[2][4];
  double sa = 0, sb = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    sa = fma(a[i].x, a[i].x, fma(a[i].y, a[i].y, sa));
    sb = fma(b[i].x, b[i].x, fma(b[i].y, b[i].y, sb));
  }
  for (int off = 32; off > 0; off >>= 1) {
    sa += __shfl_down(sa, off, 64);
    sb += __shfl_down(sb, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = sa;
    sh[1][threadIdx.x >> 6] = sb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[2 * (long long)blockIdx.x] = make_double2(sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3], 0.0);
    partial[2 * (long long)blockIdx.x + 1] = make_double2(sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3], 0.0);
  }
}

// One FP32 pass of the banded overlap factor along one rank, out-of-place:
// MODE 0: y = U x, MODE 1: y = U^T x (U real upper band, [j][sd + i - j]).
template <int SD, int MODE>
__global__ __launch_bounds__(256) void overlap_elem32_kernel(const float* __restrict__ u, int n, long long post,
                                                             long long total, const float2* __restrict__ x,
                                                             float2* __restrict__ y) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int i = (int)((idx / post) % n);
  const long long base = idx - (long long)i * post;
  float2 acc = make_float2(0.f, 0.f);
#pragma unroll
  for (int d = 0; d <= SD; ++d) {
    const int j = MODE == 0 ? i + d : i - d;
    if (j >= 0 && j < n) {
      const float a = MODE == 0 ? u[j * (SD + 1) + SD - d] : u[i * (SD + 1) + SD - d];
      const float2 v = x[base + (long long)j * post];
      acc.x = fmaf(a, v.x, acc.x);
      acc.y = fmaf(a, v.y, acc.y);
    }
  }
  y[idx] = acc;
}

// S1 (x) S2 on one channel per workgroup: the four passes U1, U1^T (along r1), U2, U2^T
// (along r2) of the per-pass kernel above, with the same per-element sums, on an n1 x n2
// block held in LDS with both factors.  Rank 0 (the channel) has no overlap.  Only one
// workgroup per channel runs, so each has 16 waves and reads only LDS after its loads.
template <int SD>
__global__ __launch_bounds__(1024) void overlap_ss32_kernel(const float* __restrict__ gu1, int n1,
                                                            const float* __restrict__ gu2, int n2,
                                                            const float2* __restrict__ x, float2* __restrict__ y) {
  extern __shared__ float2 ssbuf[];  // two n1 x n2 blocks, then both factors
  const int m = n1 * n2;
  float2* A = ssbuf;
  float2* B = ssbuf + m;
  float* u1 = reinterpret_cast<float*>(ssbuf + 2 * m);
  float* u2 = u1 + n1 * (SD + 1);
  const long long off = (long long)blockIdx.x * m;
  for (int e = threadIdx.x; e < m; e += blockDim.x) A[e] = x[off + e];
  for (int e = threadIdx.x; e < n1 * (SD + 1); e += blockDim.x) u1[e] = gu1[e];
  for (int e = threadIdx.x; e < n2 * (SD + 1); e += blockDim.x) u2[e] = gu2[e];
  __syncthreads();
  for (int e = threadIdx.x; e < m; e += blockDim.x) {  // B = U1 A
    const int i = e / n2, c = e - i * n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i + d;
      if (j < n1) {
        const float a = u1[j * (SD + 1) + SD - d];
        const float2 v = A[j * n2 + c];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    B[e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m; e += blockDim.x) {  // A = U1^T B
    const int i = e / n2, c = e - i * n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i - d;
      if (j >= 0) {
        const float a = u1[i * (SD + 1) + SD - d];
        const float2 v = B[j * n2 + c];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    A[e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m; e += blockDim.x) {  // B = U2 A (along the contiguous axis)
    const int i1 = e / n2, i = e - i1 * n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i + d;
      if (j < n2) {
        const float a = u2[j * (SD + 1) + SD - d];
        const float2 v = A[i1 * n2 + j];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    B[e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m; e += blockDim.x) {  // y = U2^T B
    const int i1 = e / n2, i = e - i1 * n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i - d;
      if (j >= 0) {
        const float a = u2[i * (SD + 1) + SD - d];
        const float2 v = B[i1 * n2 + j];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    y[off + e] = acc;
  }
}

constexpr int kSSLoads = 8;  // input elements per thread staged by overlap_ss32_rows_kernel (x 256)

// The same four passes with a channel spread over ceil(n1 / RB) workgroups: workgroup
// (channel, block) owns r1 rows [r0, r0 + RB).  U1 couples row i to rows i..i+SD and U1^T
// to rows i-SD..i, so the r1 passes of the owned rows need input rows r0-SD .. r0+RB+SD-1
// (recomputed by the neighbours, no exchange); the r2 passes act within a row.  Each output
// element is summed in the same order as in overlap_ss32_kernel, so the results are equal.
template <int SD, int RB>
__global__ __launch_bounds__(256) void overlap_ss32_rows_kernel(const float* __restrict__ gu1, int n1,
                                                                const float* __restrict__ gu2, int n2,
                                                                const float2* __restrict__ x, float2* __restrict__ y) {
  extern __shared__ float2 ssbuf[];
  constexpr int HR = RB + 2 * SD;  // rows held: owned rows and both halos
  float2* A = ssbuf;               // input rows a0 .. a1-1 at A[(i - a0) * n2 + c]
  float2* B = ssbuf + HR * n2;     // first pass, then the r2 passes
  float* u1 = reinterpret_cast<float*>(ssbuf + 2 * HR * n2);
  float* u2 = u1 + n1 * (SD + 1);
  const int c = blockIdx.x, r0 = blockIdx.y * RB, r1 = min(n1, r0 + RB);
  const int a0 = max(0, r0 - SD), a1 = min(n1, r1 + SD);  // input rows needed
  const int b0 = a0, b1 = r1;                             // first-pass rows needed
  const long long off = (long long)c * n1 * n2;
  // all loads of a thread issued before its LDS stores (the host guarantees that the input
  // rows fit in kSSLoads x 256 elements and each factor in 4 x 256)
  const int na = (a1 - a0) * n2;
  float2 av[kSSLoads];
#pragma unroll
  for (int q = 0; q < kSSLoads; ++q) {
    const int e = threadIdx.x + q * 256;
    av[q] = e < na ? x[off + (long long)a0 * n2 + e] : make_float2(0.f, 0.f);
  }
  float f1[4], f2[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + q * 256;
    f1[q] = e < n1 * (SD + 1) ? gu1[e] : 0.f;
    f2[q] = e < n2 * (SD + 1) ? gu2[e] : 0.f;
  }
#pragma unroll
  for (int q = 0; q < kSSLoads; ++q) {
    const int e = threadIdx.x + q * 256;
    if (e < na) A[e] = av[q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + q * 256;
    if (e < n1 * (SD + 1)) u1[e] = f1[q];
    if (e < n2 * (SD + 1)) u2[e] = f2[q];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < (b1 - b0) * n2; e += blockDim.x) {  // B = U1 A, rows b0 .. b1-1
    const int i = b0 + e / n2, cc = e % n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i + d;
      if (j < n1) {
        const float a = u1[j * (SD + 1) + SD - d];
        const float2 v = A[(j - a0) * n2 + cc];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    B[(i - b0) * n2 + cc] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < (r1 - r0) * n2; e += blockDim.x) {  // A = U1^T B, owned rows
    const int i = r0 + e / n2, cc = e % n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i - d;
      if (j >= 0) {
        const float a = u1[i * (SD + 1) + SD - d];
        const float2 v = B[(j - b0) * n2 + cc];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    A[(i - r0) * n2 + cc] = acc;  // owned row i now at A[i - r0] (rows read above stay in B)
  }
  __syncthreads();
  for (int e = threadIdx.x; e < (r1 - r0) * n2; e += blockDim.x) {  // B = U2 A along r2
    const int i1 = e / n2, i = e - i1 * n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i + d;
      if (j < n2) {
        const float a = u2[j * (SD + 1) + SD - d];
        const float2 v = A[i1 * n2 + j];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    B[e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < (r1 - r0) * n2; e += blockDim.x) {  // y = U2^T B
    const int i1 = e / n2, i = e - i1 * n2;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int d = 0; d <= SD; ++d) {
      const int j = i - d;
      if (j >= 0) {
        const float a = u2[i * (SD + 1) + SD - d];
        const float2 v = B[i1 * n2 + j];
        acc.x = fmaf(a, v.x, acc.x);
        acc.y = fmaf(a, v.y, acc.y);
      }
    }
    y[off + (long long)(r0 + i1) * n2 + i] = acc;
  }
}

constexpr int kSSRows = 8;  // r1 rows per workgroup of overlap_ss32_rows_kernel

void MixedWork::overlap_multiply32(const Repr& r, const float2* in, float2* out, float2* scratch, hipStream_t s) {
  for (int rank = 0; rank < r.ndim; ++rank) {
    const OverlapFactor* ov = r.ov[rank];
    if (ov && u32[rank].count == 0) {
      std::vector<float> h(ov->u_host.begin(), ov->u_host.end());
      u32[rank].upload(h.data(), h.size(), s);
    }
  }
  // (channel, r1, r2) with both radial overlaps: one launch, one channel per workgroup
  if (r.ndim == 3 && !r.ov[0] && r.ov[1] && r.ov[2] && r.ov[1]->sd == r.ov[2]->sd &&
      r.ov[1]->n == r.extents[1] && r.ov[2]->n == r.extents[2] &&
      2 * (size_t)r.extents[1] * r.extents[2] * sizeof(float2) +
              (size_t)(r.extents[1] + r.extents[2]) * (r.ov[1]->sd + 1) * sizeof(float) <=
          65536 &&
      in != out) {
    const int n1 = r.extents[1], n2 = r.extents[2];
    const int sd = r.ov[1]->sd;
    const size_t factors = (size_t)(n1 + n2) * (sd + 1) * sizeof(float);
    const size_t lds_rows = 2 * (size_t)(kSSRows + 2 * sd) * n2 * sizeof(float2) + factors;
    if (!ss_one_wg && lds_rows <= 65536 && (kSSRows + 2 * sd) * n2 <= kSSLoads * 256 &&
        (size_t)std::max(n1, n2) * (sd + 1) <= 4 * 256) {  // several workgroups per channel
      const dim3 grid(r.extents[0], ceil_div(n1, kSSRows));
      switch (sd) {
#define PP_SSR(K)                                                                                          \
  case K:                                                                                                  \
    hipLaunchKernelGGL((overlap_ss32_rows_kernel<K, kSSRows>), grid, dim3(256), lds_rows, s, u32[1].ptr, n1, \
                       u32[2].ptr, n2, in, out);                                                           \
    PP_HIP(hipGetLastError());                                                                             \
    return;
        PP_SSR(1) PP_SSR(2) PP_SSR(3) PP_SSR(4) PP_SSR(5) PP_SSR(6) PP_SSR(7) PP_SSR(8)
#undef PP_SSR
        default:
          break;
      }
    }
    const size_t lds = 2 * (size_t)n1 * n2 * sizeof(float2) + factors;
    switch (sd) {
#define PP_SS(K)                                                                                       \
  case K:                                                                                              \
    hipLaunchKernelGGL((overlap_ss32_kernel<K>), dim3(r.extents[0]), dim3(1024), lds, s, u32[1].ptr, n1, \
                       u32[2].ptr, n2, in, out);                                                       \
    PP_HIP(hipGetLastError());                                                                         \
    return;
      PP_SS(1) PP_SS(2) PP_SS(3) PP_SS(4) PP_SS(5) PP_SS(6) PP_SS(7) PP_SS(8)
#undef PP_SS
      default:
        break;
    }
  }
  const float2* cur = in;
  bool any = false;
  for (int rank = 0; rank < r.ndim; ++rank) {
    const OverlapFactor* ov = r.ov[rank];
    if (!ov) continue;
    long long post = 1;
    for (int q = rank + 1; q < r.ndim; ++q) post *= r.extents[q];
    const long long total = n;
    const int eb = (int)ceil_div(total, 256);
    auto pass = [&](int mode, const float2* x, float2* y) {
      switch (ov->sd * 2 + mode) {
#define PP_EL(K)                                                                                                      \
  case 2 * K:                                                                                                         \
    hipLaunchKernelGGL((overlap_elem32_kernel<K, 0>), dim3(eb), dim3(256), 0, s, u32[rank].ptr, ov->n, post, total, x, y); \
    break;                                                                                                            \
  case 2 * K + 1:                                                                                                     \
    hipLaunchKernelGGL((overlap_elem32_kernel<K, 1>), dim3(eb), dim3(256), 0, s, u32[rank].ptr, ov->n, post, total, x, y); \
    break;
        PP_EL(1) PP_EL(2) PP_EL(3) PP_EL(4) PP_EL(5) PP_EL(6) PP_EL(7) PP_EL(8)
#undef PP_EL
        default:
          throw Error(PP_ERR_INVALID, "FP32 overlap supports 1..8 superdiagonals");
      }
    };
    pass(0, cur, scratch);
    pass(1, scratch, out);
    cur = out;
    any = true;
  }
  if (!any && in != out) PP_HIP(hipMemcpyAsync(out, in, n * sizeof(float2), hipMemcpyDeviceToDevice, s));
  PP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------
// Kronecker preconditioner
// ------------------------------------------------------------------------------------
// One stage of K_c^-1 on an n x n complex block held in LDS, 2 x 2 outputs per thread:
//   LEFT:  out(i, j) = sum_k m(i, k) in(k, j)
//   RIGHT: out(i, j) = sum_k in(i, k) m(k, j)
// m(p, q) = M[p*n + q], or M[q*n + p] when TRANS.  SCALE 1 divides by, SCALE 2 multiplies
// by, d_ij = 1 + a (lam1_i + lam2_j).
template <class R, class C, bool LEFT, bool TRANS, int SCALE>
__device__ inline void kron_stage(const R* __restrict__ M, const C* in, C* out, int n, const R* lam1,
                                  const R* lam2, C a) {
  const int h = (n + 1) >> 1;
  for (int t = threadIdx.x; t < h * h; t += blockDim.x) {
    const int i0 = (t / h) * 2, j0 = (t % h) * 2;
    const bool i1ok = i0 + 1 < n, j1ok = j0 + 1 < n;
    C acc[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        acc[p][q].x = 0;
        acc[p][q].y = 0;
      }
    for (int k = 0; k < n; ++k) {
      if (LEFT) {
        // m(i, k) for i = i0, i0+1 ; in(k, j) for j = j0, j0+1
        const R m0 = TRANS ? M[k * n + i0] : M[i0 * n + k];
        const R m1 = i1ok ? (TRANS ? M[k * n + i0 + 1] : M[(i0 + 1) * n + k]) : R(0);
        const C v0 = in[k * n + j0];
        const C v1 = j1ok ? in[k * n + j0 + 1] : v0;
        acc[0][0].x = fma(m0, v0.x, acc[0][0].x); acc[0][0].y = fma(m0, v0.y, acc[0][0].y);
        acc[0][1].x = fma(m0, v1.x, acc[0][1].x); acc[0][1].y = fma(m0, v1.y, acc[0][1].y);
        acc[1][0].x = fma(m1, v0.x, acc[1][0].x); acc[1][0].y = fma(m1, v0.y, acc[1][0].y);
        acc[1][1].x = fma(m1, v1.x, acc[1][1].x); acc[1][1].y = fma(m1, v1.y, acc[1][1].y);
      } else {
        // in(i, k) for i = i0, i0+1 ; m(k, j) for j = j0, j0+1
        const C v0 = in[i0 * n + k];
        const C v1 = i1ok ? in[(i0 + 1) * n + k] : v0;
        const R m0 = TRANS ? M[j0 * n + k] : M[k * n + j0];
        const R m1 = j1ok ? (TRANS ? M[(j0 + 1) * n + k] : M[k * n + j0 + 1]) : R(0);
        acc[0][0].x = fma(m0, v0.x, acc[0][0].x); acc[0][0].y = fma(m0, v0.y, acc[0][0].y);
        acc[0][1].x = fma(m1, v0.x, acc[0][1].x); acc[0][1].y = fma(m1, v0.y, acc[0][1].y);
        acc[1][0].x = fma(m0, v1.x, acc[1][0].x); acc[1][0].y = fma(m0, v1.y, acc[1][0].y);
        acc[1][1].x = fma(m1, v1.x, acc[1][1].x); acc[1][1].y = fma(m1, v1.y, acc[1][1].y);
      }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int i = i0 + p, j = j0 + q;
        if (i >= n || j >= n) continue;
        C v = acc[p][q];
        if (SCALE == 1) {
          const R sl = lam1[i] + lam2[j];
          const R dr = R(1) + a.x * sl, di = a.y * sl;
          const R inv = R(1) / (dr * dr + di * di);
          C qv;
          qv.x = (v.x * dr + v.y * di) * inv;
          qv.y = (v.y * dr - v.x * di) * inv;
          v = qv;
        } else if (SCALE == 2) {
          const R sl = lam1[i] + lam2[j];
          const R dr = R(1) + a.x * sl, di = a.y * sl;
          C qv;
          qv.x = v.x * dr - v.y * di;
          qv.y = v.x * di + v.y * dr;
          v = qv;
        }
        out[i * n + j] = v;
      }
  }
}

// One workgroup per channel c.  E_IN_LDS stages both eigenvector matrices in LDS (FP32);
// the FP64 variant reads them through the cache to stay inside 160 KiB of LDS.
// FWD applies K_c itself: with F_l = S E_l (= E_l^-T),  K_c B = F_l1 [(F_l1^T B F_l2) d] F_l2^T,
// E then holds F.
template <class R, class C, bool E_IN_LDS, bool FWD>
__global__ __launch_bounds__(1024) void kron_kernel(const R* __restrict__ E, const R* __restrict__ lam,
                                                    const int* __restrict__ l1, const int* __restrict__ l2, int n,
                                                    C a, const C* in, C* psi, R in_scale, C* in_wb, const C* add,
                                                    const double* add_norm, R add_sigma, const double2* in_n2) {
  if (in_n2) {  // scale = 1 / |in| from a squared norm on the device (0 for a zero vector)
    const double q = in_n2->x;
    in_scale = q > 0 ? (R)(1.0 / sqrt(q)) : R(0);
  }
  extern __shared__ unsigned char smem[];
  const int c = blockIdx.x;
  const int n2 = n * n;
  C* B = reinterpret_cast<C*>(smem);
  C* T = B + n2;
  const R* E1 = E + (long long)l1[c] * n2;
  const R* E2 = E + (long long)l2[c] * n2;
  if (E_IN_LDS) {
    R* e1 = reinterpret_cast<R*>(T + n2);
    R* e2 = e1 + n2;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      e1[i] = E1[i];
      e2[i] = E2[i];
    }
    E1 = e1;
    E2 = e2;
  }
  const R* lam1 = lam + (long long)l1[c] * n;
  const R* lam2 = lam + (long long)l2[c] * n;
  C* x = psi + (long long)c * n2;
  const C* xin = in + (long long)c * n2;  // may alias x: read completely before any write
  const R as = add ? add_sigma * (R)*add_norm : R(0);
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    C v = xin[i];
    v.x *= in_scale;
    v.y *= in_scale;
    if (add) {  // in + sigma |v| add (the learned correction), |v| read on the device
      v.x += as * add[(long long)c * n2 + i].x;
      v.y += as * add[(long long)c * n2 + i].y;
    }
    B[i] = v;
    if (in_wb) in_wb[(long long)c * n2 + i] = v;  // the scaled input, in place of in
  }
  __syncthreads();
  kron_stage<R, C, true, true, 0>(E1, B, T, n, lam1, lam2, a);              // T = E1^T B
  __syncthreads();
  kron_stage<R, C, false, false, FWD ? 2 : 1>(E2, T, B, n, lam1, lam2, a);  // B = (T E2) / d  (or * d)
  __syncthreads();
  kron_stage<R, C, true, false, 0>(E1, B, T, n, lam1, lam2, a);             // T = E1 B
  __syncthreads();
  kron_stage<R, C, false, true, 0>(E2, T, B, n, lam1, lam2, a);             // B = T E2^T
  __syncthreads();
  for (int i = threadIdx.x; i < n2; i += blockDim.x) x[i] = B[i];
}

// FP32 MFMA form for n <= 64: the block is padded to 64 x 64, kept as separate re/im planes
// in LDS (row stride 65 floats), and each of the 16 waves owns one 16 x 16 output tile
// (re and im accumulators) of every stage: 16 K steps of v_mfma_f32_16x16x4f32 per tile.
typedef float kf32x4 __attribute__((ext_vector_type(4)));
constexpr int KLD = 65;

// C = op(M) In (LEFT) or In op(M) (RIGHT) on 64 x 64 planes; op(M)(p, q) = TRANS ? M[q][p] : M[p][q]
template <bool LEFT, bool TRANS, int SCALE>
__device__ inline void kron_mfma_stage(const float* M, const float* in_re, const float* in_im, float* out_re,
                                       float* out_im, int n, const float* lam1, const float* lam2, float2 a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ti = wave >> 2, tj = wave & 3;
  const int lr = lane & 15, lk = lane >> 4;
  kf32x4 cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
  for (int k0 = 0; k0 < 64; k0 += 4) {
    const int k = k0 + lk;
    if (LEFT) {
      // A = op(M)[ti*16 + lr][k], B = In[k][tj*16 + lr]
      const int i = ti * 16 + lr;
      const float av = TRANS ? M[k * KLD + i] : M[i * KLD + k];
      const int j = tj * 16 + lr;
      cre = __builtin_amdgcn_mfma_f32_16x16x4f32(av, in_re[k * KLD + j], cre, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f32_16x16x4f32(av, in_im[k * KLD + j], cim, 0, 0, 0);
    } else {
      // A = In[ti*16 + lr][k], B = op(M)[k][tj*16 + lr]
      const int i = ti * 16 + lr;
      const int j = tj * 16 + lr;
      const float bv = TRANS ? M[j * KLD + k] : M[k * KLD + j];
      cre = __builtin_amdgcn_mfma_f32_16x16x4f32(in_re[i * KLD + k], bv, cre, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f32_16x16x4f32(in_im[i * KLD + k], bv, cim, 0, 0, 0);
    }
  }
  // lane holds C[4 lk + r][lr] of the tile
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ti * 16 + lk * 4 + r, j = tj * 16 + lr;
    float vr = cre[r], vi = cim[r];
    if (SCALE == 1) {
      if (i < n && j < n) {
        const float sl = lam1[i] + lam2[j];
        const float dr = 1.f + a.x * sl, di = a.y * sl;
        const float inv = 1.f / (dr * dr + di * di);
        const float qr = (vr * dr + vi * di) * inv, qi = (vi * dr - vr * di) * inv;
        vr = qr;
        vi = qi;
      } else {
        vr = 0.f;
        vi = 0.f;
      }
    }
    out_re[i * KLD + j] = vr;
    out_im[i * KLD + j] = vi;
  }
}

__global__ __launch_bounds__(1024) void kron_mfma_kernel(const float* __restrict__ E, const float* __restrict__ lam,
                                                         const int* __restrict__ l1, const int* __restrict__ l2,
                                                         int n, float2 a, const float2* in, float2* psi,
                                                         float in_scale, float2* in_wb, const float2* __restrict__ add,
                                                         const double* __restrict__ add_norm, float add_sigma,
                                                         const double2* __restrict__ in_n2) {
  if (in_n2) {  // scale = 1 / |in| from a squared norm on the device (0 for a zero vector)
    const double q = in_n2->x;
    in_scale = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
  }
  extern __shared__ float kl[];
  float* Bre = kl;
  float* Bim = Bre + 64 * KLD;
  float* Tre = Bim + 64 * KLD;
  float* Tim = Tre + 64 * KLD;
  float* e1 = Tim + 64 * KLD;
  float* e2 = e1 + 64 * KLD;
  const int c = blockIdx.x;
  const int n2 = n * n;
  const float* E1 = E + (long long)l1[c] * n2;
  const float* E2 = E + (long long)l2[c] * n2;
  const float* lam1 = lam + (long long)l1[c] * n;
  const float* lam2 = lam + (long long)l2[c] * n;
  float2* x = psi + (long long)c * n2;
  const float2* xin = in + (long long)c * n2;  // may alias x: read completely before any write
  for (int idx = threadIdx.x; idx < 64 * 64; idx += blockDim.x) {
    const int i = idx >> 6, j = idx & 63;
    const bool ok = i < n && j < n;
    float2 v = ok ? make_float2(in_scale * xin[i * n + j].x, in_scale * xin[i * n + j].y) : make_float2(0.f, 0.f);
    if (ok && add) {  // in + sigma |v| add (the learned correction), |v| read on the device
      const float as = add_sigma * (float)*add_norm;
      const float2 y = add[(long long)c * n2 + i * n + j];
      v.x = fmaf(as, y.x, v.x);
      v.y = fmaf(as, y.y, v.y);
    }
    if (ok && in_wb) in_wb[(long long)c * n2 + i * n + j] = v;  // the scaled input, in place of in
    Bre[i * KLD + j] = v.x;
    Bim[i * KLD + j] = v.y;
    e1[i * KLD + j] = ok ? E1[i * n + j] : 0.f;
    e2[i * KLD + j] = ok ? E2[i * n + j] : 0.f;
  }
  __syncthreads();
  kron_mfma_stage<true, true, 0>(e1, Bre, Bim, Tre, Tim, n, lam1, lam2, a);    // T = E1^T B
  __syncthreads();
  kron_mfma_stage<false, false, 1>(e2, Tre, Tim, Bre, Bim, n, lam1, lam2, a);  // B = (T E2) / d
  __syncthreads();
  kron_mfma_stage<true, false, 0>(e1, Bre, Bim, Tre, Tim, n, lam1, lam2, a);   // T = E1 B
  __syncthreads();
  kron_mfma_stage<false, true, 0>(e2, Tre, Tim, Bre, Bim, n, lam1, lam2, a);   // B = T E2^T
  __syncthreads();
  for (int idx = threadIdx.x; idx < n2; idx += blockDim.x) {
    const int i = idx / n, j = idx - i * n;
    x[idx] = make_float2(Bre[i * KLD + j], Bim[i * KLD + j]);
  }
}

// The same FP32 solve split over four workgroups per channel in two launches (the
// one-workgroup form keeps all four stages of a channel on one CU, where the MFMA work of a
// channel takes longer than the rest of the kernel):
//   kron_split1_kernel  workgroup (c, t): rows 16t..16t+15 of G = (E1^T B E2) / d, B the
//                       (scaled, corrected) input, via U = E1[:, strip]^T B; G to g32
//   kron_split2_kernel  workgroup (c, t): rows 16t..16t+15 of X = E1 G E2^T, via V = E1[strip, :] G
// Four waves per workgroup, wave w owns the 16-column tile w of each strip product.
// Strip product of 16 x 64 (LDS, row stride KLD) by 64 x 64 (LDS): out(i, j) = sum_k
// L(i, k) R(k, j) with L(i, k) = LT ? lm[k][i0 + i] : lm[i0 + i][k] and R(k, j) = RT ? rm[j][k] : rm[k][j]
// (j0: the first output column of the wave's 16 x 16 tile)
template <bool LT, bool RT>
__device__ inline void kron_strip(const float* lm, int i0, const float* r_re, const float* r_im, bool r_complex,
                                  const float* l_im, int j0, kf32x4& cre, kf32x4& cim) {
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int j = j0 + lr;
#pragma unroll 4
  for (int k0 = 0; k0 < 64; k0 += 4) {
    const int k = k0 + lk;
    const float lv = LT ? lm[k * KLD + i0 + lr] : lm[(i0 + lr) * KLD + k];
    const float rre = RT ? r_re[j * KLD + k] : r_re[k * KLD + j];
    if (r_complex) {  // real left factor, complex right factor
      const float rim = RT ? r_im[j * KLD + k] : r_im[k * KLD + j];
      cre = __builtin_amdgcn_mfma_f32_16x16x4f32(lv, rre, cre, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f32_16x16x4f32(lv, rim, cim, 0, 0, 0);
    } else {  // complex left factor (lm re, l_im im), real right factor
      const float liv = LT ? l_im[k * KLD + i0 + lr] : l_im[(i0 + lr) * KLD + k];
      cre = __builtin_amdgcn_mfma_f32_16x16x4f32(lv, rre, cre, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f32_16x16x4f32(liv, rre, cim, 0, 0, 0);
    }
  }
}

__global__ __launch_bounds__(1024) void kron_split1_kernel(const float* __restrict__ E, const float* __restrict__ lam,
                                                           const int* __restrict__ l1, const int* __restrict__ l2,
                                                           int n, float2 a, const float2* in, float in_scale,
                                                           const float2* __restrict__ add,
                                                           const double* __restrict__ add_norm, float add_sigma,
                                                           const double2* __restrict__ in_n2, float2* __restrict__ g) {
  if (in_n2) {
    const double q = in_n2->x;
    in_scale = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
  }
  extern __shared__ float kl[];
  float* Bre = kl;                // 64 x KLD
  float* Bim = Bre + 64 * KLD;
  float* e1 = Bim + 64 * KLD;     // columns i0 .. i0+15 only
  float* e2 = e1 + 64 * KLD;
  float* Ure = e2 + 64 * KLD;     // 16 x KLD
  float* Uim = Ure + 16 * KLD;
  const int c = blockIdx.x, t = blockIdx.y, tid = threadIdx.x;
  const int i0 = t * 16;
  const int n2 = n * n;
  const float* E1 = E + (long long)l1[c] * n2;
  const float* E2 = E + (long long)l2[c] * n2;
  const float2* xin = in + (long long)c * n2;
  // every load of a thread is issued before the LDS stores: 4 of B, 4 of E2, 1 of the E1 strip
  float2 bv[4];
  float ev2[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * 1024, i = idx >> 6, j = idx & 63;
    const bool ok = i < n && j < n;
    bv[q] = ok ? xin[i * n + j] : make_float2(0.f, 0.f);
    ev2[q] = ok ? E2[i * n + j] : 0.f;
  }
  const int sk = tid >> 4, si = i0 + (tid & 15);
  const float ev1 = (sk < n && si < n) ? E1[sk * n + si] : 0.f;
  const float as = add ? add_sigma * (float)*add_norm : 0.f;
  float2 yv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * 1024, i = idx >> 6, j = idx & 63;
    yv[q] = (add && i < n && j < n) ? add[(long long)c * n2 + i * n + j] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * 1024, i = idx >> 6, j = idx & 63;
    float2 v = make_float2(in_scale * bv[q].x, in_scale * bv[q].y);
    if (add) {  // in + sigma |v| add (the learned correction)
      v.x = fmaf(as, yv[q].x, v.x);
      v.y = fmaf(as, yv[q].y, v.y);
    }
    Bre[i * KLD + j] = v.x;
    Bim[i * KLD + j] = v.y;
    e2[i * KLD + j] = ev2[q];
  }
  e1[sk * KLD + si] = ev1;
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  kf32x4 cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
  if (wave < 4) {
    kron_strip<true, false>(e1, i0, Bre, Bim, true, nullptr, wave * 16, cre, cim);  // U = E1[:, strip]^T B
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = (lane >> 4) * 4 + r, j = wave * 16 + (lane & 15);
      Ure[i * KLD + j] = cre[r];
      Uim[i * KLD + j] = cim[r];
    }
  }
  __syncthreads();
  if (wave >= 4) return;
  cre = kf32x4{0.f, 0.f, 0.f, 0.f};
  cim = kf32x4{0.f, 0.f, 0.f, 0.f};
  kron_strip<false, false>(Ure, 0, e2, nullptr, false, Uim, wave * 16, cre, cim);  // G strip = U E2
  const float* lam1 = lam + (long long)l1[c] * n;
  const float* lam2 = lam + (long long)l2[c] * n;
  float2* gc = g + (long long)c * 64 * 64;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + (lane >> 4) * 4 + r, j = wave * 16 + (lane & 15);
    float vr = 0.f, vi = 0.f;
    if (i < n && j < n) {  // divide by d = 1 + a (lam1_i + lam2_j), as kron_mfma_stage
      const float sl = lam1[i] + lam2[j];
      const float dr = 1.f + a.x * sl, di = a.y * sl;
      const float inv = 1.f / (dr * dr + di * di);
      vr = (cre[r] * dr + cim[r] * di) * inv;
      vi = (cim[r] * dr - cre[r] * di) * inv;
    }
    gc[i * 64 + j] = make_float2(vr, vi);
  }
}

__global__ __launch_bounds__(1024) void kron_split2_kernel(const float* __restrict__ E, const int* __restrict__ l1,
                                                           const int* __restrict__ l2, int n,
                                                           const float2* __restrict__ g, const float2* in, float2* out,
                                                           float2* in_wb, float in_scale,
                                                           const double2* __restrict__ in_n2) {
  extern __shared__ float kl[];
  float* Gre = kl;                // 64 x KLD
  float* Gim = Gre + 64 * KLD;
  float* e1 = Gim + 64 * KLD;     // rows i0 .. i0+15 only
  float* e2 = e1 + 64 * KLD;
  float* Vre = e2 + 64 * KLD;     // 16 x KLD
  float* Vim = Vre + 16 * KLD;
  const int c = blockIdx.x, t = blockIdx.y, tid = threadIdx.x;
  const int i0 = t * 16;
  const int n2 = n * n;
  const float* E1 = E + (long long)l1[c] * n2;
  const float* E2 = E + (long long)l2[c] * n2;
  const float2* gc = g + (long long)c * 64 * 64;
  float2 gv[4];
  float ev2[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * 1024, i = idx >> 6, j = idx & 63;
    const bool ok = i < n && j < n;  // rows past the last strip are never written: read as 0
    gv[q] = ok ? gc[idx] : make_float2(0.f, 0.f);
    ev2[q] = ok ? E2[i * n + j] : 0.f;
  }
  const int si = i0 + (tid >> 6), sk = tid & 63;
  const float ev1 = (si < n && sk < n) ? E1[si * n + sk] : 0.f;
  float2* xb = in_wb ? in_wb + (long long)c * n2 : nullptr;
  float2 wv = make_float2(0.f, 0.f);
  const int wi = i0 + tid / 64, wj = tid % 64;  // this thread's element of the write-back strip
  const bool wok = xb && wi < n && wj < n && wi < i0 + 16;
  if (wok) wv = in[(long long)c * n2 + wi * n + wj];  // the input (the write-back may be elsewhere)
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * 1024, i = idx >> 6, j = idx & 63;
    Gre[i * KLD + j] = gv[q].x;
    Gim[i * KLD + j] = gv[q].y;
    e2[i * KLD + j] = ev2[q];
  }
  e1[si * KLD + sk] = ev1;
  if (wok) {  // the scaled input, in place of in (rows of this strip; in does not alias out here)
    if (in_n2) {
      const double q = in_n2->x;
      in_scale = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
    }
    xb[wi * n + wj] = make_float2(in_scale * wv.x, in_scale * wv.y);
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  kf32x4 cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
  if (wave < 4) {
    kron_strip<false, false>(e1, i0, Gre, Gim, true, nullptr, wave * 16, cre, cim);  // V = E1[strip, :] G
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = (lane >> 4) * 4 + r, j = wave * 16 + (lane & 15);
      Vre[i * KLD + j] = cre[r];
      Vim[i * KLD + j] = cim[r];
    }
  }
  __syncthreads();
  if (wave >= 4) return;
  cre = kf32x4{0.f, 0.f, 0.f, 0.f};
  cim = kf32x4{0.f, 0.f, 0.f, 0.f};
  kron_strip<false, true>(Vre, 0, e2, nullptr, false, Vim, wave * 16, cre, cim);  // X strip = V E2^T
  float2* x = out + (long long)c * n2;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + (lane >> 4) * 4 + r, j = wave * 16 + (lane & 15);
    if (i < n && j < n) x[i * n + j] = make_float2(cre[r], cim[r]);
  }
}

// The same FP32 solve in one launch, workgroup (c, t) again owning rows 16t .. 16t + 15 of
// X: each workgroup forms the whole G = (E1^T B E2) / d itself (all 16 waves, one 16 x 16
// tile each: the four workgroups of a channel repeat that half of the work, which costs
// less than a second launch and the exchange of G through memory), then its strip
// V = E1[strip, :] G and X = V E2^T (waves 0-3).  Every tile is the split kernels' tile with
// the same operands and order, so the result is theirs bit for bit.  With a pending
// Gram-Schmidt finish (fin.p3), every workgroup sums the |w|^2 block partials itself (the
// sum finish_cgs1_kernel forms, the same bits) for the scale, and workgroup (0, 0) also sums
// the dots and writes the finished results in that kernel's layout: the finish takes no
// launch of its own.  With sz (and F, F_l = S E_l), waves 4-7 form the strip of the overlap
// product S1 X S2 = F1 G F2^T beside X (E^T S E = I), from the same G: V' = F1[strip, :] G,
// then V' F2^T.
__global__ __launch_bounds__(1024) void kron_fused_kernel(const float* __restrict__ E, const float* __restrict__ lam,
                                                          const int* __restrict__ l1, const int* __restrict__ l2,
                                                          int n, float2 a, const float2* in, float in_scale,
                                                          const float2* __restrict__ add,
                                                          const double* __restrict__ add_norm, float add_sigma,
                                                          const double2* __restrict__ in_n2, float2* out,
                                                          float2* in_wb, GsFinish fin, const float* __restrict__ F,
                                                          float2* sz) {
  __shared__ double2 fres[130];
  if (!fin.p3 && in_n2) {
    const double q = in_n2->x;
    in_scale = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
  }
  extern __shared__ float kl[];
  float* Bre = kl;                // B, then G (64 x KLD each plane)
  float* Bim = Bre + 64 * KLD;
  float* e1 = Bim + 64 * KLD;
  float* e2 = e1 + 64 * KLD;
  float* Ure = e2 + 64 * KLD;     // U = E1^T B
  float* Uim = Ure + 64 * KLD;
  float* Vre = Uim + 64 * KLD;    // V strip (16 x KLD)
  float* Vim = Vre + 16 * KLD;
  float* f1 = Vim + 16 * KLD;     // with sz: rows i0 .. i0+15 of F1 (16 x KLD), F2 (64 x KLD), V' strip
  float* f2 = f1 + 16 * KLD;
  float* Pre = f2 + 64 * KLD;
  float* Pim = Pre + 16 * KLD;
  const int c = blockIdx.x, t = blockIdx.y, tid = threadIdx.x;
  const int i0 = t * 16;
  const int n2 = n * n;
  const float* E1 = E + (long long)l1[c] * n2;
  const float* E2 = E + (long long)l2[c] * n2;
  const float2* xin = in + (long long)c * n2;
  // every load of a thread is issued before the LDS stores: 4 each of B, E1, E2 (and the
  // correction; with sz 4 of F2 and 1 of the F1 strip)
  float2 bv[4], yv[4];
  float ev1[4], ev2[4], fv2[4], fv1 = 0.f;
  const float* F1 = sz ? F + (long long)l1[c] * n2 : nullptr;
  const float* F2 = sz ? F + (long long)l2[c] * n2 : nullptr;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * 1024, i = idx >> 6, j = idx & 63;
    const bool ok = i < n && j < n;
    bv[q] = ok ? xin[i * n + j] : make_float2(0.f, 0.f);
    ev1[q] = ok ? E1[i * n + j] : 0.f;
    ev2[q] = ok ? E2[i * n + j] : 0.f;
    yv[q] = (add && ok) ? add[(long long)c * n2 + i * n + j] : make_float2(0.f, 0.f);
    fv2[q] = (sz && ok) ? F2[i * n + j] : 0.f;
  }
  const int fi = i0 + (tid >> 6), fk = tid & 63;  // this thread's element of the F1 strip
  if (sz && fi < n && fk < n) fv1 = F1[fi * n + fk];
  if (fin.p3) {  // after the loads above are issued, so that their latencies overlap
    const bool first = blockIdx.x == 0 && blockIdx.y == 0;
    if (first) sum_partials(fin.p1, fin.blocks, fin.m + 1, fin.m + 1, fres);
    sum_partials(fin.p3, fin.blocks, 1, 1, fres + fin.m + 1);
    __syncthreads();
    const double q = fres[fin.m + 1].x;
    in_scale = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
    if (first) {
      const int width = fin.m + 1;
      for (int j = tid; j < 2 * width + 1; j += blockDim.x) {
        const double2 v = j < width ? fres[j] : (j == 2 * width ? fres[width] : make_double2(0.0, 0.0));
        fin.out[j] = v;
        if (fin.out2) fin.out2[j] = v;
      }
    }
  }
  const float as = add ? add_sigma * (float)*add_norm : 0.f;
  float2* xb = in_wb ? in_wb + (long long)c * n2 : nullptr;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * 1024, i = idx >> 6, j = idx & 63;
    const float2 sv = make_float2(in_scale * bv[q].x, in_scale * bv[q].y);
    float2 v = sv;
    if (add) {  // in + sigma |v| add (the learned correction)
      v.x = fmaf(as, yv[q].x, v.x);
      v.y = fmaf(as, yv[q].y, v.y);
    }
    Bre[i * KLD + j] = v.x;
    Bim[i * KLD + j] = v.y;
    e1[i * KLD + j] = ev1[q];
    e2[i * KLD + j] = ev2[q];
    if (sz) f2[i * KLD + j] = fv2[q];
    // the scaled input in place of in, rows of this strip (in does not alias out here)
    if (xb && i >= i0 && i < i0 + 16 && i < n && j < n) xb[i * n + j] = sv;
  }
  if (sz) f1[(tid >> 6) * KLD + fk] = fv1;
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  const int si = wave >> 2, ct = wave & 3;  // this wave's tile of a 64 x 64 product
  kf32x4 cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
  kron_strip<true, false>(e1, si * 16, Bre, Bim, true, nullptr, ct * 16, cre, cim);  // U = E1^T B
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = si * 16 + (lane >> 4) * 4 + r, j = ct * 16 + (lane & 15);
    Ure[i * KLD + j] = cre[r];
    Uim[i * KLD + j] = cim[r];
  }
  __syncthreads();
  cre = kf32x4{0.f, 0.f, 0.f, 0.f};
  cim = kf32x4{0.f, 0.f, 0.f, 0.f};
  kron_strip<false, false>(Ure, si * 16, e2, nullptr, false, Uim, ct * 16, cre, cim);  // U E2
  const float* lam1 = lam + (long long)l1[c] * n;
  const float* lam2 = lam + (long long)l2[c] * n;
#pragma unroll
  for (int r = 0; r < 4; ++r) {  // G = U E2 / d, into B's planes (B is no longer read)
    const int i = si * 16 + (lane >> 4) * 4 + r, j = ct * 16 + (lane & 15);
    float vr = 0.f, vi = 0.f;
    if (i < n && j < n) {
      const float sl = lam1[i] + lam2[j];
      const float dr = 1.f + a.x * sl, di = a.y * sl;
      const float inv = 1.f / (dr * dr + di * di);
      vr = (cre[r] * dr + cim[r] * di) * inv;
      vi = (cim[r] * dr - cre[r] * di) * inv;
    }
    Bre[i * KLD + j] = vr;
    Bim[i * KLD + j] = vi;
  }
  __syncthreads();
  if (wave < 4) {
    cre = kf32x4{0.f, 0.f, 0.f, 0.f};
    cim = kf32x4{0.f, 0.f, 0.f, 0.f};
    kron_strip<false, false>(e1, i0, Bre, Bim, true, nullptr, wave * 16, cre, cim);  // V = E1[strip, :] G
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = (lane >> 4) * 4 + r, j = wave * 16 + (lane & 15);
      Vre[i * KLD + j] = cre[r];
      Vim[i * KLD + j] = cim[r];
    }
  } else if (sz && wave < 8) {
    cre = kf32x4{0.f, 0.f, 0.f, 0.f};
    cim = kf32x4{0.f, 0.f, 0.f, 0.f};
    kron_strip<false, false>(f1, 0, Bre, Bim, true, nullptr, (wave - 4) * 16, cre, cim);  // V' = F1[strip, :] G
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = (lane >> 4) * 4 + r, j = (wave - 4) * 16 + (lane & 15);
      Pre[i * KLD + j] = cre[r];
      Pim[i * KLD + j] = cim[r];
    }
  }
  __syncthreads();
  if (wave < 4) {
    cre = kf32x4{0.f, 0.f, 0.f, 0.f};
    cim = kf32x4{0.f, 0.f, 0.f, 0.f};
    kron_strip<false, true>(Vre, 0, e2, nullptr, false, Vim, wave * 16, cre, cim);  // X strip = V E2^T
    float2* x = out + (long long)c * n2;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) * 4 + r, j = wave * 16 + (lane & 15);
      if (i < n && j < n) x[i * n + j] = make_float2(cre[r], cim[r]);
    }
  } else if (sz && wave < 8) {
    cre = kf32x4{0.f, 0.f, 0.f, 0.f};
    cim = kf32x4{0.f, 0.f, 0.f, 0.f};
    kron_strip<false, true>(Pre, 0, f2, nullptr, false, Pim, (wave - 4) * 16, cre, cim);  // strip of V' F2^T
    float2* y = sz + (long long)c * n2;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) * 4 + r, j = (wave - 4) * 16 + (lane & 15);
      if (i < n && j < n) y[i * n + j] = make_float2(cre[r], cim[r]);
    }
  }
}

// the one-launch form's LDS; with the overlap product, the F1 strip, F2 and V' as well
static size_t kron_lds_fused(bool with_sz = true) {
  return ((6 * 64 + 2 * 16) + (with_sz ? 16 + 64 + 2 * 16 : 0)) * (size_t)KLD * sizeof(float);
}

static size_t kron_lds_split() { return (4 * 64 + 2 * 16) * (size_t)KLD * sizeof(float); }

static size_t kron_lds_mfma() { return 6 * 64 * (size_t)KLD * sizeof(float); }

static size_t kron_lds64(int n) { return 2 * (size_t)n * n * sizeof(double2); }
static size_t kron_lds32(int n) { return 2 * (size_t)n * n * (sizeof(float2) + sizeof(float)); }

void Kron::setup_attributes() {
  PP_REQUIRE(kron_lds64(n) <= 160 * 1024 && kron_lds32(n) <= 160 * 1024, "Kronecker block too large for LDS");
  PP_HIP(hipFuncSetAttribute((const void*)kron_kernel<double, double2, false, false>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kron_lds64(n)));
  PP_HIP(hipFuncSetAttribute((const void*)kron_kernel<double, double2, false, true>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kron_lds64(n)));
  PP_HIP(hipFuncSetAttribute((const void*)kron_kernel<float, float2, true, false>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kron_lds32(n)));
  PP_HIP(hipFuncSetAttribute((const void*)kron_mfma_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kron_lds_mfma()));
  PP_HIP(hipFuncSetAttribute((const void*)kron_fused_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kron_lds_fused()));
  PP_HIP(hipFuncSetAttribute((const void*)kron_split1_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kron_lds_split()));
  PP_HIP(hipFuncSetAttribute((const void*)kron_split2_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kron_lds_split()));
}

void kron_solve(const Kron& k, double2 a, double2* psi, hipStream_t s) { k.apply64(a, psi, s); }

void Kron::apply64(double2 a, double2* psi, hipStream_t s) const {
  hipLaunchKernelGGL((kron_kernel<double, double2, false, false>), dim3(nchan), dim3(1024), kron_lds64(n), s, E64.ptr,
                     lam64.ptr, l1.ptr, l2.ptr, n, a, psi, psi, 1.0, (double2*)nullptr, (const double2*)nullptr,
                     (const double*)nullptr, 0.0, (const double2*)nullptr);
  PP_HIP(hipGetLastError());
}

void Kron::multiply64(double2 a, double2* psi, hipStream_t s) const {
  hipLaunchKernelGGL((kron_kernel<double, double2, false, true>), dim3(nchan), dim3(1024), kron_lds64(n), s, F64.ptr,
                     lam64.ptr, l1.ptr, l2.ptr, n, a, psi, psi, 1.0, (double2*)nullptr, (const double2*)nullptr,
                     (const double*)nullptr, 0.0, (const double2*)nullptr);
  PP_HIP(hipGetLastError());
}

void Kron::apply32(double2 a, float2* psi, hipStream_t s) const { apply32(a, psi, psi, s, 1.f, nullptr, nullptr); }

void Kron::apply32(double2 a, const float2* in, float2* out, hipStream_t s, float in_scale, float2* wb,
                   const double2* in_n2, const GsFinish* fin, float2* sz) const {
  launch32(a, in, out, s, in_scale, wb, nullptr, nullptr, 0.f, in_n2, fin, sz);
}

bool Kron::overlap_product32(const float2* in, const float2* out, const float2* wb) const {
  return n <= 64 && split32 && fused32 && F32.count > 0 && in != out && (!wb || (wb != out && wb != in));
}

void Kron::apply32_add(double2 a, const float2* in, const float2* add, const double* add_norm, float add_sigma,
                       float2* out, hipStream_t s, float2* sz) const {
  launch32(a, in, out, s, 1.f, nullptr, add, add_norm, add_sigma, nullptr, nullptr, sz);
}

void Kron::launch32(double2 a, const float2* in, float2* out, hipStream_t s, float in_scale, float2* wb,
                    const float2* add, const double* add_norm, float add_sigma, const double2* in_n2,
                    const GsFinish* fin, float2* sz) const {
  PP_REQUIRE(!fin || (fin->p3 && in_n2), "a pending finish needs the device slot of its |w|^2");
  PP_REQUIRE(!sz || overlap_product32(in, out, wb), "the overlap product needs the one-launch form");
  // one launch only without aliasing: the workgroups of a channel all read the whole input
  // block while each writes its strip of out (and of the write-back)
  if (n <= 64 && split32 && fused32 && in != out && (!wb || (wb != out && wb != in && !add))) {
    const dim3 grid(nchan, (n + 15) / 16);
    hipLaunchKernelGGL(kron_fused_kernel, grid, dim3(1024), kron_lds_fused(sz != nullptr), s, E32.ptr, lam32.ptr,
                       l1.ptr, l2.ptr, n, make_float2((float)a.x, (float)a.y), in, in_scale, add, add_norm, add_sigma,
                       in_n2, out, wb, fin ? *fin : GsFinish{}, F32.ptr, sz);
    PP_HIP(hipGetLastError());
    return;
  }
  if (fin) {  // the other forms read the finished |w|^2 from in_n2
    hipLaunchKernelGGL(finish_cgs1_kernel, dim3(1), dim3(256), 0, s, fin->blocks, fin->p1, fin->m, fin->p3, fin->out,
                       fin->out2);
    PP_HIP(hipGetLastError());
  }
  if (n <= 64 && split32 && (!wb || (wb != out && !add))) {
    if (g32.count < (size_t)nchan * 64 * 64) g32.resize((size_t)nchan * 64 * 64);
    const dim3 grid(nchan, (n + 15) / 16);
    const float2 af = make_float2((float)a.x, (float)a.y);
    hipLaunchKernelGGL(kron_split1_kernel, grid, dim3(1024), kron_lds_split(), s, E32.ptr, lam32.ptr, l1.ptr,
                       l2.ptr, n, af, in, in_scale, add, add_norm, add_sigma, in_n2, g32.ptr);
    hipLaunchKernelGGL(kron_split2_kernel, grid, dim3(1024), kron_lds_split(), s, E32.ptr, l1.ptr, l2.ptr, n,
                       g32.ptr, in, out, wb, in_scale, in_n2);
    PP_HIP(hipGetLastError());
    return;
  }
  if (n <= 64) {
    hipLaunchKernelGGL(kron_mfma_kernel, dim3(nchan), dim3(1024), kron_lds_mfma(), s, E32.ptr, lam32.ptr, l1.ptr,
                       l2.ptr, n, make_float2((float)a.x, (float)a.y), in, out, in_scale, wb, add, add_norm,
                       add_sigma, in_n2);
    PP_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL((kron_kernel<float, float2, true, false>), dim3(nchan), dim3(1024), kron_lds32(n), s, E32.ptr,
                     lam32.ptr, l1.ptr, l2.ptr, n, make_float2((float)a.x, (float)a.y), in, out, in_scale, wb, add,
                     add_norm, add_sigma, in_n2);
  PP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------
// flexible GMRES in FP32
// ------------------------------------------------------------------------------------
// min ||A y - b|| by Householder QR (A column-major rows x cols, full column rank)
static void least_squares_small(int rows, int cols, std::vector<std::complex<double>> A,
                                std::vector<std::complex<double>> b, std::vector<std::complex<double>>& x) {
  using cd = std::complex<double>;
  for (int k = 0; k < cols; ++k) {
    double nrm = 0;
    for (int i = k; i < rows; ++i) nrm += std::norm(A[(size_t)k * rows + i]);
    nrm = std::sqrt(nrm);
    if (nrm == 0) continue;
    const cd akk = A[(size_t)k * rows + k];
    const cd alpha = std::abs(akk) == 0 ? cd(-nrm, 0) : -akk / std::abs(akk) * nrm;
    std::vector<cd> v(rows - k);
    v[0] = akk - alpha;
    for (int i = k + 1; i < rows; ++i) v[i - k] = A[(size_t)k * rows + i];
    double vn = 0;
    for (auto& z : v) vn += std::norm(z);
    if (vn == 0) continue;
    for (int c = k; c < cols; ++c) {
      cd d = 0;
      for (int i = k; i < rows; ++i) d += std::conj(v[i - k]) * A[(size_t)c * rows + i];
      d *= 2.0 / vn;
      for (int i = k; i < rows; ++i) A[(size_t)c * rows + i] -= d * v[i - k];
    }
    cd d = 0;
    for (int i = k; i < rows; ++i) d += std::conj(v[i - k]) * b[i];
    d *= 2.0 / vn;
    for (int i = k; i < rows; ++i) b[i] -= d * v[i - k];
  }
  x.assign(cols, 0);
  for (int k = cols - 1; k >= 0; --k) {
    cd sum = b[k];
    for (int c = k + 1; c < cols; ++c) sum -= A[(size_t)c * rows + k] * x[c];
    x[k] = A[(size_t)k * rows + k] == cd(0) ? cd(0) : sum / A[(size_t)k * rows + k];
  }
}

void MixedWork::ensure(long long n_, int m_) {
  if (n == n_ && m >= m_) return;
  n = n_;
  m = m_;
  V.resize((size_t)(m + 1) * n);
  Z.resize((size_t)m * n);
  w.resize(n);
  tmp.resize(n);
  wnext.resize(n);
  sz.resize(n);
  sz_for = nullptr;
  if (!small) {
    // results the host reads: kernels write them into mapped, coherent pinned memory, so a
    // read is one stream synchronisation and no copy
    PP_HIP(hipHostMalloc((void**)&small, kSmall * sizeof(double2), hipHostMallocMapped | hipHostMallocCoherent));
    PP_HIP(hipHostGetDevicePointer((void**)&small_dev, small, 0));
  }
  // per inner iteration j: 2 (j + 2) + 1 results at stride 2 (m + 1) + 1, then 8 step-level slots
  PP_REQUIRE((long long)m * (2 * (m + 1) + 1) <= kSmall - 8, "inner basis too large for the host result block");
  b64.resize(n);
  r64.resize(n);
  x64.resize(n);
  t64.resize(n);
  h64.resize(n);
  norms.resize(4);
  hdev.resize(kSmall);
  ycoef.resize(kDevLsMax);
}

void MixedWork::read_small(int count, hipStream_t s) {
  PP_REQUIRE(count <= kSmall, "host result block too small");
  PP_HIP(hipStreamSynchronize(s));  // the finishing kernels wrote straight into small
  host_small.assign(small, small + count);
}

MixedWork::~MixedWork() {
  if (small) (void)hipHostFree(small);
}

void MixedWork::norm2_pair(const double2* a, const double2* b, double2* out, double2* out_dev, hipStream_t s) {
  const int blocks = fused_blocks(n);
  red.ensure(blocks, 2);
  hipLaunchKernelGGL(norm2_pair_partial_kernel, dim3(blocks), dim3(256), 0, s, n, a, b, red.partial.ptr);
  hipLaunchKernelGGL(finish_partials_kernel, dim3(1), dim3(256), 0, s, blocks, red.partial.ptr, 2, 2, nullptr, 0, 0,
                     nullptr, 0, 0, out, out_dev);
  PP_HIP(hipGetLastError());
}

// Classical Gram-Schmidt, two passes, against V_0..V_{k-1} in four launches:
// out[0..k] = first dots (+|w|^2), out[k+1..2k+1] = second dots (+|w|^2), out[2k+2] = |w|^2
// after both updates.
void MixedWork::cgs2_32(int k, float2* w_, double2* out, double2* out_dev, hipStream_t s) {
  const int blocks = fused_blocks(n);
  const int width = k + 1;
  // the finishing kernel holds all 2 width + 1 results in a 128-entry LDS array
  PP_REQUIRE(2 * width + 1 <= 128, "Gram-Schmidt basis too large");
  red.ensure(blocks, 2 * width + 1);
  double2* p1 = red.partial.ptr;
  double2* p2 = p1 + (size_t)blocks * width;
  double2* p3 = p2 + (size_t)blocks * width;
  const dim3 g(blocks), b(256);
  hipLaunchKernelGGL((cgs32_kernel<false, true>), g, b, 0, s, n, k, V.ptr, n, nullptr, 0, w_, p1);
  hipLaunchKernelGGL((cgs32_kernel<true, true>), g, b, 0, s, n, k, V.ptr, n, p1, width, w_, p2);
  hipLaunchKernelGGL((cgs32_kernel<true, false>), g, b, 0, s, n, k, V.ptr, n, p2, width, w_, p3, wnext.ptr);
  hipLaunchKernelGGL(finish_partials_kernel, dim3(1), b, 0, s, blocks, p1, width, width, p2, width, width, p3, 1, 1,
                     out, out_dev);
  PP_HIP(hipGetLastError());
}

// One-pass classical Gram-Schmidt against V_0..V_{k-1} in three launches: the dots (+|w|^2),
// the update with the norm of the updated vector, and a one-block finish (results in the
// two-pass layout, see finish_cgs1_kernel)
void MixedWork::cgs1_32(int k, float2* w_, double2* out, double2* out_dev, hipStream_t s, bool defer) {
  PP_REQUIRE(!pend.p3, "a Gram-Schmidt finish is still pending");
  const int blocks = fused_blocks(n);
  const int width = k + 1;
  PP_REQUIRE(2 * width + 1 <= 128, "Gram-Schmidt basis too large");
  red.ensure(blocks, width + 1);
  double2* p1 = red.partial.ptr;
  double2* p3 = p1 + (size_t)blocks * width;
  const dim3 g(blocks), b(256);
  hipLaunchKernelGGL((cgs32_kernel<false, true>), g, b, 0, s, n, k, V.ptr, n, nullptr, 0, w_, p1);
  hipLaunchKernelGGL((cgs32_kernel<true, false>), g, b, 0, s, n, k, V.ptr, n, p1, width, w_, p3, wnext.ptr);
  if (defer) {
    pend.p1 = p1;
    pend.p3 = p3;
    pend.blocks = blocks;
    pend.m = k;
    pend.out = out;
    pend.out2 = out_dev;
  } else {
    hipLaunchKernelGGL(finish_cgs1_kernel, dim3(1), b, 0, s, blocks, p1, k, p3, out, out_dev);
  }
  PP_HIP(hipGetLastError());
}

void MixedWork::flush_finish(hipStream_t s) {
  if (!pend.p3) return;
  hipLaunchKernelGGL(finish_cgs1_kernel, dim3(1), dim3(256), 0, s, pend.blocks, pend.p1, pend.m, pend.p3, pend.out,
                     pend.out2);
  PP_HIP(hipGetLastError());
  pend = GsFinish{};
}

void MixedWork::gram_schmidt(int k, float2* w_, double2* out, double2* out_dev, hipStream_t s, bool defer) {
  if (gs_passes == 1) cgs1_32(k, w_, out, out_dev, s, defer && defer_finish);
  else cgs2_32(k, w_, out, out_dev, s);
}

// out[0] = |w|^2 (w is only read)
void MixedWork::norm2_32(const float2* w_, double2* out, hipStream_t s) {
  const int blocks = fused_blocks(n);
  red.ensure(blocks, 1);
  hipLaunchKernelGGL((cgs32_kernel<false, false>), dim3(blocks), dim3(256), 0, s, n, 0, V.ptr, n, nullptr, 0,
                     const_cast<float2*>(w_), red.partial.ptr);
  hipLaunchKernelGGL(finish_partials_kernel, dim3(1), dim3(256), 0, s, blocks, red.partial.ptr, 1, 1, nullptr, 0, 0,
                     nullptr, 0, 0, out, (double2*)nullptr);
  PP_HIP(hipGetLastError());
}

void MixedWork::finish_norms(long long blocks, int count, double2* out, double2* out_dev, hipStream_t s) {
  PP_REQUIRE(blocks >= 1 && blocks < (1LL << 31) && count >= 1 && count <= 2, "invalid norm partials");
  hipLaunchKernelGGL(finish_norms_kernel, dim3(1), dim3(256), 0, s, (int)blocks,
                     reinterpret_cast<const double2*>(npart.ptr), count, out, out_dev);
  PP_HIP(hipGetLastError());
}

void MixedWork::residual(const double2* b64_, double2* r, double2* out, double2* out_dev, hipStream_t s) {
  const int blocks = fused_blocks(n);
  red.ensure(blocks, 1);
  hipLaunchKernelGGL(residual_partial_kernel, dim3(blocks), dim3(256), 0, s, n, b64_, r, red.partial.ptr);
  hipLaunchKernelGGL(finish_partials_kernel, dim3(1), dim3(256), 0, s, blocks, red.partial.ptr, 1, 1, nullptr, 0, 0,
                     nullptr, 0, 0, out, out_dev);
  PP_HIP(hipGetLastError());
}

int MixedWork::fgmres32(const std::function<void(const float2*, float2*)>& apply_A,
                        const std::function<void(int, const float2*, const double2*, float2*, float2*)>& apply_P, const float2* b,
                        double rtol, int max_it, double* rel_out, hipStream_t s, double beta_in,
                        const double2* base,
                        double2* xout, const std::function<bool(double&)>& first_read, int pre_iters,
                        int* needed) {
  using cd = std::complex<double>;
  const int mm = std::min(max_it, m);
  PP_REQUIRE(mm >= 1 && mm + 3 <= 120, "inner basis size out of range");
  double beta = beta_in;
  if (b) {
    norm2_32(b, small_dev, s);
    read_small(1, s);
    beta = std::sqrt(host_small[0].x);
  }
  auto no_solve = [&]() {
    if (base != xout) PP_HIP(hipMemcpyAsync(xout, base, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    if (rel_out) *rel_out = 0;
    if (needed) *needed = 0;
    return 0;
  };
  if (!first_read && !(beta > 0)) return no_solve();
  if (b) hipLaunchKernelGGL(scale32_kernel, dim3(grid_n(n)), dim3(256), 0, s, n, (float)(1.0 / beta), b, V.ptr);
  std::vector<cd> H((size_t)(mm + 1) * mm, 0.0);  // column-major (mm+1) x mm
  auto Hc = [&](int row, int col) -> cd& { return H[(size_t)col * (mm + 1) + row]; };
  std::vector<cd> y;
  int k = 0;
  double rel = 1.0;
  // Results of iteration j go to their own slot of the pinned block (stride) and of its
  // device copy.  The Arnoldi process needs no host value: iteration j normalises V_j by
  // the device copy of |w|^2 from iteration j - 1.  So the first `pre` iterations run
  // before the host reads anything; after that the host reads every iteration.
  const int stride = 2 * (mm + 1) + 1;
  PP_REQUIRE(mm * stride <= kSmall - 8, "inner basis too large for the host result block");
  int pre = std::max(1, std::min(pre_iters, mm));
  int processed = 0, first_conv = 0;
  bool done = false;
  for (int j = 0; j < mm && !done; ++j) {
    float2* vj = V.ptr + (size_t)j * n;
    float2* zj = Z.ptr + (size_t)j * n;
    float2* vn = V.ptr + (size_t)(j + 1) * n;
    const double2* n2 = j == 0 ? nullptr : hdev.ptr + (size_t)(j - 1) * stride + 2 * (j + 1);
    // V_0 is already normalised; later directions come as the updated vector in wnext and
    // are normalised into V_j by the preconditioner's load
    apply_P(j, j == 0 ? vj : wnext.ptr, n2, zj, j == 0 ? nullptr : vj);
    apply_A(zj, vn);
    // classical Gram-Schmidt: [first dots, |w|^2 | second dots, |w|^2 | final |w|^2]; the
    // updated vector goes to wnext.  Before the host reads, the next preconditioner launch
    // finishes it
    gram_schmidt(j + 1, vn, small_dev + (size_t)j * stride, hdev.ptr + (size_t)j * stride, s, j + 1 < pre);
    if (j + 1 < pre) continue;
    PP_HIP(hipStreamSynchronize(s));  // the finishing kernels wrote straight into small
    for (int jj = processed; jj <= j && !done; ++jj) {
      const double2* hs = small + (size_t)jj * stride;
      const int w1 = jj + 2;  // h (jj+1) + |w|^2
      if (jj == 0 && first_read) {
        // the caller learns |rhs| (and may decide the solve is not needed) only now
        if (!first_read(beta)) {
          if (rel_out) *rel_out = 0;
          if (needed) *needed = 0;
          return -1;
        }
        if (!(beta > 0)) return no_solve();
      }
      for (int i = 0; i <= jj; ++i) Hc(i, jj) = cd(hs[i].x + hs[w1 + i].x, hs[i].y + hs[w1 + i].y);
      const double hn = std::sqrt(std::max(0.0, hs[2 * w1].x));
      Hc(jj + 1, jj) = hn;
      k = jj + 1;
      std::vector<cd> A((size_t)(k + 1) * k), rhs_ls(k + 1, 0.0);
      for (int c = 0; c < k; ++c)
        for (int r = 0; r <= k; ++r) A[(size_t)c * (k + 1) + r] = Hc(r, c);
      rhs_ls[0] = beta;
      least_squares_small(k + 1, k, A, rhs_ls, y);
      double e2 = 0;
      for (int r = 0; r <= k; ++r) {
        cd acc = r == 0 ? -beta : 0.0;
        for (int c = 0; c < k; ++c) acc += Hc(r, c) * y[c];
        e2 += std::norm(acc);
      }
      rel = std::sqrt(e2) / beta;
      processed = jj + 1;
      if (rel <= rtol && !first_conv) first_conv = k;
      // a breakdown ends the basis here (later columns, if any were run, are not used);
      // otherwise every column already computed is used, since each lowers the residual
      if (!(hn > 1e-30 * beta)) done = true;
    }
    if (first_conv) done = true;
    pre = 0;
  }
  flush_finish(s);
  if (needed) *needed = first_conv ? first_conv : k;
  LincombCoef yc;
  for (int i = 0; i < k; ++i) yc.c[i] = make_float2((float)y[i].real(), (float)y[i].imag());
  hipLaunchKernelGGL(lincomb_add32_kernel, dim3(grid_n(n)), dim3(256), 0, s, n, k, Z.ptr, n, yc, base, xout);
  PP_HIP(hipGetLastError());
  if (rel_out) *rel_out = rel;
  return k;
}

// k inner iterations with no host read: iteration j normalises V_j from the device copy of
// iteration j-1's |w|^2, the least-squares problem is solved on the device and
// xout = base + Z y is formed from the device coefficients.  V_0 must hold the normalised
// right-hand side, whose squared norm is beta2_dev->x.  The residual estimate after each
// iteration goes to small[est .. est + k), readable after the caller's next synchronisation.
void MixedWork::fgmres32_device(const std::function<void(const float2*, float2*)>& apply_A,
                                const std::function<void(int, const float2*, const double2*, float2*, float2*)>& apply_P, int k,
                                const double2* beta2_dev, const double2* base, double2* xout, int est,
                                hipStream_t s) {
  PP_REQUIRE(k >= 1 && k <= kDevLsMax && k <= m, "device inner solve size out of range");
  PP_REQUIRE(est >= 0 && est + k <= kSmall, "estimate slots outside the host result block");
  const int stride = 2 * (k + 1) + 1;
  for (int j = 0; j < k; ++j) {
    float2* vj = V.ptr + (size_t)j * n;
    float2* zj = Z.ptr + (size_t)j * n;
    float2* vn = V.ptr + (size_t)(j + 1) * n;
    const double2* n2 = j == 0 ? nullptr : hdev.ptr + (size_t)(j - 1) * stride + 2 * (j + 1);
    apply_P(j, j == 0 ? vj : wnext.ptr, n2, zj, j == 0 ? nullptr : vj);
    apply_A(zj, vn);
    // all but the last finished by the next preconditioner launch
    gram_schmidt(j + 1, vn, small_dev + (size_t)j * stride, hdev.ptr + (size_t)j * stride, s, j + 1 < k);
  }
  flush_finish(s);
  hipLaunchKernelGGL(hessenberg_lsq_kernel, dim3(1), dim3(64), 0, s, k, stride, hdev.ptr, beta2_dev, ycoef.ptr,
                     small_dev + est);
  hipLaunchKernelGGL(lincomb_add32_dev_kernel, dim3(grid_n(n)), dim3(256), 0, s, n, k, Z.ptr, n, ycoef.ptr, base,
                     xout);
  PP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------
// mixed Cayley step
// ------------------------------------------------------------------------------------
void MixedCayley::step(double2* psi, double dt, const double* sc, hipStream_t s) {
  PP_REQUIRE(ham && kron, "mixed Cayley needs a Hamiltonian and a Kronecker preconditioner");
  Hamiltonian& H = *ham;
  const Repr& R = *H.repr;
  const long long n = H.n;
  work.ensure(n, m);
  scal.assign(sc, sc + H.terms.size());
  stats = MixedStats{};
  const double a = 0.5 * dt;
  const Kron& K = *kron;
  const double2 ka = make_double2(0.0, a);
  // with an overlap term in the factored operator, the FP64 S + c H is one multiply
  const bool fused_s = fham && fham->has_overlap;
  // FP64 sparse angular form: its epilogue writes the block partials of the outputs' norms
  const long long nblk64 = fused_s ? fham->norm_blocks64() : 0;
  if (nblk64 > 0 && work.npart.count < (size_t)(2 * nblk64)) work.npart.resize((size_t)(2 * nblk64));
  // In graph mode the step's time scalings also go to device memory (through pinned host
  // memory, in stream order): the sparse angular form reads them there, so that a captured
  // step replays with each step's values.  Eager steps pass them as kernel arguments (the
  // copy is a launch of its own).
  const double* scd = nullptr;
  if (fham && fham->sparse_angular() && graph_mode) {
    PP_REQUIRE(scal.size() <= (size_t)kMaxFTerms, "too many potentials");
    if (!scal_host) PP_HIP(hipHostMalloc((void**)&scal_host, kMaxFTerms * sizeof(double), hipHostMallocDefault));
    std::copy(scal.begin(), scal.end(), scal_host);
    if (scal_dev.count < scal.size()) scal_dev.resize(scal.size());
    PP_HIP(hipMemcpyAsync(scal_dev.ptr, scal_host, scal.size() * sizeof(double), hipMemcpyHostToDevice, s));
    scd = scal_dev.ptr;
  }
  hipStream_t cs = s;  // the stream the step's work is issued on (the capture stream while capturing)
  // |b|^2 and |r|^2: written to the pinned block for the host and to the device for the
  // scaling of V_0; the host reads them at the first synchronisation of the first correction
  constexpr int kB = MixedWork::kSmall - 4, kR0 = MixedWork::kSmall - 3, kR = MixedWork::kSmall - 2;
  constexpr int kEst = MixedWork::kSmall - 8 - 8 * kDevLsMax;  // device-mode residual estimates
  // b = S psi - i a H psi ; r = b - A psi = -2 i a H psi ; x = psi
  const bool reuse = fused_s && z_valid && z_psi == psi && fham->gen64 == z_gen;
  auto issue_start = [&]() {
    if (fused_s) {
      // Both from one multiply.  The last residual of the previous step formed the radial
      // products (overlap term included) at the accepted solution; when psi still holds it --
      // compared on the device -- only the angular stage with the new time scalings runs.
      FHamOut ob, orr;
      ob.dst = work.b64.ptr;
      ob.alpha = make_double2(0.0, -a);
      ob.beta = make_double2(1.0, 0.0);
      orr.dst = work.r64.ptr;
      orr.alpha = make_double2(0.0, -2.0 * a);
      orr.beta = make_double2(0.0, 0.0);
      // with the sparse angular form its epilogue also writes the block partials of |b|^2
      // and |r|^2, so no separate norm pass reads the two vectors again
      fham->apply64_ex(scal.data(), psi, ob, &orr, reuse ? work.x64.ptr : nullptr, cs, scd,
                       nblk64 > 0 ? work.npart.ptr : nullptr);
      if (nblk64 > 0) {
        ++stats.outer_ops;
        work.finish_norms(nblk64, 2, work.small_dev + kB, work.norms.ptr, cs);
        return;
      }
    } else {
      vec_copy(n, psi, work.t64.ptr, cs);
      R.overlap(PP_OVERLAP_MULTIPLY, false, work.t64.ptr, cs);
      if (fham) fham->apply64(scal.data(), psi, work.h64.ptr, nullptr, make_double2(1.0, 0.0), cs,
                              make_double2(0.0, 0.0), scd);
      else H.apply(scal.data(), psi, work.h64.ptr, nullptr, make_double2(1.0, 0.0), cs);
      vec_combine(n, make_double2(1.0, 0.0), work.t64.ptr, make_double2(0.0, -a), work.h64.ptr, work.b64.ptr, cs);
      vec_combine(n, make_double2(0.0, 0.0), work.h64.ptr, make_double2(0.0, -2.0 * a), work.h64.ptr,
                  work.r64.ptr, cs);
    }
    ++stats.outer_ops;
    work.norm2_pair(work.b64.ptr, work.r64.ptr, work.small_dev + kB, work.norms.ptr, cs);
  };

  auto apply_A32 = [&](const float2* in, float2* out) {
    // out = S in + i a H in.  S stays in its factored form U^T U here even when the factored
    // operator holds S (x) S: in FP32, S (x) S applied entry by entry loses its small
    // eigenvalues (relative error ~ eps32 cond(S (x) S)), and the refinement then cannot
    // reduce the error in those directions, where the residual barely sees it.
    if (fham && fham->csd == fham->bw && fham->sep_enabled) {
      // S in Cholesky form inside the multiply, or (S (x) S) in = F1 G F2^T from the
      // Kronecker solve that produced in (accurate in the same small-eigenvalue directions)
      const float2* zs = work.sz_for == in ? work.sz.ptr : nullptr;
      work.sz_for = nullptr;
      fham->apply32(scal.data(), in, out, nullptr, make_double2(0.0, a), cs, make_double2(1.0, 0.0), scd, zs);
    } else {
      work.overlap_multiply32(R, in, work.tmp.ptr, work.w.ptr, cs);
      if (fham) fham->apply32(scal.data(), in, out, work.tmp.ptr, make_double2(0.0, a), cs, make_double2(0.0, 0.0),
                              scd);
      else H.apply_f32(scal.data(), in, out, work.tmp.ptr, make_float2(0.f, (float)a), cs);
    }
    ++stats.inner_ops;
  };
  int correction = 0;
  // the solves form the FP32 multiply's overlap product (Cholesky-form overlap only)
  const bool sz_from_solve = overlap_from_solve && fham && fham->csd == fham->bw && fham->sep_enabled;
  // z = P (in / sqrt(n2)) (n2 = NULL: in is already normalised); wb (with n2): receives
  // in / sqrt(n2), the normalised basis vector
  auto apply_P32 = [&](int j, const float2* in, const double2* n2, float2* out, float2* wb) {
    // learned_mode 0: the first direction of the first correction, whose right-hand side
    // r = -2 i a H psi is the kind of vector the network was trained on; 1: the first
    // direction of every correction; 2: every direction
    const bool use = learned && (learned_mode == 2 || (j == 0 && (learned_mode == 1 || correction == 0)));
    if (use) {
      work.flush_finish(cs);  // the learned path reads the finished |w|^2
      const float2* v = in;
      if (n2) {  // V_j is normalised here (the Kronecker path does it while loading)
        hipLaunchKernelGGL(unit32_to_kernel, dim3(grid_n(n)), dim3(256), 0, cs, n, n2, in, wb);
        PP_HIP(hipGetLastError());
        v = wb;
      }
      const bool with_sz = sz_from_solve && K.overlap_product32(v, out, nullptr);
      learned_apply(*learned, K, ka, v, out, cs, with_sz ? work.sz.ptr : nullptr);
      work.sz_for = with_sz ? out : nullptr;
      ++stats.learned_applies;
    } else {
      // a finish left by the last Gram-Schmidt is done by this launch; with the Cholesky-form
      // FP32 overlap, the solve also forms (S (x) S) out for the multiply that follows
      float2* wb2 = n2 ? wb : nullptr;
      const bool with_sz = sz_from_solve && K.overlap_product32(in, out, wb2);
      K.apply32(ka, in, out, cs, 1.f, wb2, n2, work.pend.p3 ? &work.pend : nullptr, with_sz ? work.sz.ptr : nullptr);
      work.pend = GsFinish{};
      work.sz_for = with_sz ? out : nullptr;
    }
    ++stats.precond_applies;
  };
  bool converged = false;
  bool x_is_psi = true;  // x has not been formed yet: it equals psi
  // r = b - (S x + i a H x) into r64 (the residual pass follows)
  auto residual_multiply = [&]() {
    if (fused_s) {
      fham->apply64(scal.data(), work.x64.ptr, work.r64.ptr, nullptr, make_double2(0.0, a), cs,
                    make_double2(1.0, 0.0), scd);
    } else {
      vec_copy(n, work.x64.ptr, work.t64.ptr, cs);
      R.overlap(PP_OVERLAP_MULTIPLY, false, work.t64.ptr, cs);
      if (fham) fham->apply64(scal.data(), work.x64.ptr, work.r64.ptr, work.t64.ptr, make_double2(0.0, a), cs,
                              make_double2(0.0, 0.0), scd);
      else H.apply(scal.data(), work.x64.ptr, work.r64.ptr, work.t64.ptr, make_double2(0.0, a), cs);
    }
    ++stats.outer_ops;
  };
  // r = b - A x with |r|^2 finished in the same launches: the FP64 angular epilogue forms
  // b - (S x + i a H x) and its block partials (sparse angular form), else the multiply and a
  // residual pass
  auto residual_pass = [&](hipStream_t st) {
    if (fused_s && nblk64 > 0) {
      FHamOut o;
      o.dst = work.r64.ptr;
      o.addend = work.b64.ptr;
      o.alpha = make_double2(0.0, -a);
      o.beta = make_double2(-1.0, 0.0);
      fham->apply64_ex(scal.data(), work.x64.ptr, o, nullptr, nullptr, cs, scd, work.npart.ptr);
      ++stats.outer_ops;
      work.finish_norms(nblk64, 1, work.small_dev + kR, work.norms.ptr + 2, st);
    } else {
      residual_multiply();
      work.residual(work.b64.ptr, work.r64.ptr, work.small_dev + kR, work.norms.ptr + 2, st);
    }
  };
  double bnorm = -1.0, rn = -1.0;
  // a residual norm written by a kernel and not yet read: its index in the pinned block and
  // its device copy
  int pending = kR0;
  const double2* pending_dev = work.norms.ptr + 1;
  double predicted = HUGE_VAL;  // relative residual the pending value is expected to show
  auto read_pending = [&]() {   // the stream has been synchronised
    if (bnorm < 0) bnorm = std::sqrt(std::max(0.0, work.small[kB].x));
    rn = std::sqrt(std::max(0.0, work.small[pending].x));
    pending = -1;
    stats.residual = rn;
    stats.rel_residual = bnorm > 0 ? rn / bnorm : 0.0;
    return rn <= tol_rel * bnorm;
  };
  // Device mode: when the previous step has shown how many corrections and inner iterations
  // a step needs, run them with no host read in between (each inner least-squares problem
  // solved on the device) and synchronise once, at the final residual.  A step that is not
  // converged then continues below, one host-driven correction at a time.
  int dev_corr = device_mode && prev_corrections > 0 && (int)pre_its.size() >= prev_corrections ? prev_corrections
                                                                                                  : 0;
  for (int c = 0; c < dev_corr; ++c)
    if (pre_its[c] > kDevLsMax || pre_its[c] > m) dev_corr = 0;
  auto issue_device = [&]() {
    for (int c = 0; c < dev_corr; ++c) {
      ++stats.outer_iterations;
      hipLaunchKernelGGL(f64_to_f32_unit_kernel, dim3(grid_n(n)), dim3(256), 0, cs, n, pending_dev, work.r64.ptr,
                         work.V.ptr);
      work.fgmres32_device(apply_A32, apply_P32, pre_its[c], pending_dev, x_is_psi ? psi : work.x64.ptr,
                           work.x64.ptr, kEst + c * kDevLsMax, cs);
      stats.inner_iterations += pre_its[c];
      x_is_psi = false;
      ++correction;
      residual_pass(cs);
      pending = kR;
      pending_dev = work.norms.ptr + 2;
    }
  };
  // Graph mode: a device-mode step is one fixed sequence of launches for given counts, so it
  // is captured once (on an internal stream, in relaxed mode) and replayed on s; the time
  // scalings come from device memory.  A sequence is captured only after it has run
  // eagerly with the same counts (every buffer it uses has then reached its size, so the
  // capture allocates and clears nothing).
  std::vector<long long> key;
  if (dev_corr > 0) {
    key = {dev_corr, reuse ? 1 : 0, (long long)(intptr_t)psi, learned ? learned_mode : -1};
    for (int c = 0; c < dev_corr; ++c) key.push_back(pre_its[c]);
  }
  const bool use_graph = dev_corr > 0 && graph_mode && graph_ok && scd &&
                         (!reuse || fham->reuse_flag.count > 0) && ((gexec && key == gkey) || key == eager_key);
  if (use_graph) {
    if (!gexec || key != gkey) {
      if (!cap_stream) PP_HIP(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
      cs = cap_stream;
      PP_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
      hipGraph_t g = nullptr;
      try {
        issue_start();
        issue_device();
      } catch (...) {
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        cs = s;
        graph_ok = false;
        throw;
      }
      const hipError_t e = hipStreamEndCapture(cs, &g);
      cs = s;
      if (e != hipSuccess || !g) {
        graph_ok = false;
        throw Error(PP_ERR_HIP, std::string("capture of the mixed step failed: ") + hipGetErrorString(e));
      }
      if (gexec) (void)hipGraphExecDestroy(gexec);
      gexec = nullptr;
      const hipError_t ei = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        gexec = nullptr;
        graph_ok = false;
        throw Error(PP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      }
      gkey = key;
      gstats = stats;  // what issuing the sequence counts, applied again on every replay
      ++graph_captures;
    } else {
      stats = gstats;
      correction = dev_corr;
      x_is_psi = false;
      pending = kR;
      pending_dev = work.norms.ptr + 2;
    }
    PP_HIP(hipGraphLaunch(gexec, s));
    ++graph_launches;
  } else {
    issue_start();
    if (dev_corr > 0) {
      issue_device();
      eager_key = key;
    }
  }
  z_valid = false;
  if (dev_corr > 0) {
    PP_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < dev_corr; ++c) {
      // the next step runs as many iterations as this correction needed to reach inner_rtol
      const int k = pre_its[c];
      int needed = 0;
      for (int j = 0; j < k && !needed; ++j)
        if (work.small[kEst + c * kDevLsMax + j].x <= inner_rtol) needed = j + 1;
      pre_its[c] = needed ? needed : std::min(k + 1, std::min(kDevLsMax, m));
    }
    if (read_pending()) converged = true;
    predicted = HUGE_VAL;
  }
  for (int outer = stats.outer_iterations; !converged; ++outer) {
    const bool limit = outer >= max_outer || stats.inner_iterations > fallback_iterations;
    // Read a pending residual now when it is likely converged or a limit is reached; else
    // the next correction starts at once and the residual is read at that correction's
    // first synchronisation (a needless correction is then abandoned after one iteration).
    if (pending >= 0 && (predicted <= 10.0 * tol_rel || limit)) {
      PP_HIP(hipStreamSynchronize(s));
      if (read_pending()) {
        converged = true;
        break;
      }
    }
    if (limit) break;
    ++stats.outer_iterations;
    // V_0 = r / |r| in FP32, the norm read on the device
    hipLaunchKernelGGL(f64_to_f32_unit_kernel, dim3(grid_n(n)), dim3(256), 0, s, n, pending_dev, work.r64.ptr,
                       work.V.ptr);
    double rel = 0;
    // x = (psi or x) + d, d from the inner solve
    if ((int)pre_its.size() <= correction) pre_its.resize(correction + 1, 1);
    int needed = 0;
    const int its = work.fgmres32(
        apply_A32, apply_P32, nullptr, inner_rtol, m, &rel, s, 0.0, x_is_psi ? psi : work.x64.ptr, work.x64.ptr,
        [&](double& beta) {
          if (pending >= 0 && read_pending()) return false;  // already converged: no correction
          beta = rn;
          return true;
        },
        pre_its[correction], &needed);
    // the next step runs as many iterations of this correction before its first read
    if (its >= 0) pre_its[correction] = std::max(1, needed);
    if (its < 0) {
      --stats.outer_iterations;
      converged = true;
      break;
    }
    stats.inner_iterations += its;
    x_is_psi = false;
    ++correction;
    residual_pass(s);
    pending = kR;
    pending_dev = work.norms.ptr + 2;
    predicted = (bnorm > 0 ? rn / bnorm : 0.0) * rel;
  }
  if (converged) prev_corrections = std::max(1, stats.outer_iterations);
  if (converged && fused_s && !x_is_psi) {
    // W64 now holds the radial products (overlap term included) of x64, which becomes psi below
    z_valid = true;
    z_gen = fham->gen64;
    z_psi = psi;
  }
  if (!converged) {
    if (!fallback) throw Error(PP_ERR_NOT_CONVERGED, "mixed-precision Cayley step did not converge");
    // psi is untouched: the plain FP64 solver with the exact block preconditioner redoes the step
    fallback->step(psi, dt, sc, s);
    stats.fallback = 1;
    return;
  }
  if (!x_is_psi) vec_copy(n, work.x64.ptr, psi, s);
}

MixedCayley::~MixedCayley() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (cap_stream) (void)hipStreamDestroy(cap_stream);
  if (scal_host) (void)hipHostFree(scal_host);
}

}  // namespace pp

// ------------------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------------------
using namespace pp;

extern "C" int pp_kron_create(int nchan, int n, int nl, const double* S, const double* E, const double* lam,
                              const int* l1, const int* l2, pp_kron** out) {
  return guarded([&] {
    PP_REQUIRE(out && S && E && lam && l1 && l2 && nchan > 0 && n > 0 && nl > 0, "invalid Kronecker arguments");
    for (int c = 0; c < nchan; ++c)
      PP_REQUIRE(l1[c] >= 0 && l1[c] < nl && l2[c] >= 0 && l2[c] < nl, "channel l index out of range");
    auto* p = new pp_kron();
    Kron& k = p->k;
    k.nchan = nchan;
    k.n = n;
    k.nl = nl;
    try {
      k.setup_attributes();
      hipStream_t s = nullptr;
      const size_t ne = (size_t)nl * n * n, nlam = (size_t)nl * n;
      k.E64.upload(E, ne, s);
      k.lam64.upload(lam, nlam, s);
      std::vector<double> F(ne, 0.0);  // F_l = S E_l
      for (int l = 0; l < nl; ++l)
        for (int i = 0; i < n; ++i)
          for (int q = 0; q < n; ++q) {
            const double sv = S[(size_t)i * n + q];
            if (sv == 0.0) continue;
            for (int j = 0; j < n; ++j) F[((size_t)l * n + i) * n + j] += sv * E[((size_t)l * n + q) * n + j];
          }
      k.F64.upload(F.data(), ne, s);
      std::vector<float> ff(F.begin(), F.end());
      k.F32.upload(ff.data(), ne, s);
      std::vector<float> ef(E, E + ne), lf(lam, lam + nlam);
      k.E32.upload(ef.data(), ne, s);
      k.lam32.upload(lf.data(), nlam, s);
      k.l1.upload(l1, nchan, s);
      k.l2.upload(l2, nchan, s);
      PP_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

extern "C" int pp_kron_destroy(pp_kron* k) {
  return guarded([&] { delete k; });
}

extern "C" int pp_kron_solve(const pp_kron* k, pp_cplx a, pp_cplx* psi, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(k && psi, "null argument");
    k->k.apply64(make_double2(a.re, a.im), reinterpret_cast<double2*>(psi), (hipStream_t)stream);
  });
}

extern "C" int pp_kron_multiply(const pp_kron* k, pp_cplx a, pp_cplx* psi, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(k && psi, "null argument");
    k->k.multiply64(make_double2(a.re, a.im), reinterpret_cast<double2*>(psi), (hipStream_t)stream);
  });
}

extern "C" int pp_kron_solve_f32(const pp_kron* k, pp_cplx a, pp_cplxf* psi, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(k && psi, "null argument");
    k->k.apply32(make_double2(a.re, a.im), reinterpret_cast<float2*>(psi), (hipStream_t)stream);
  });
}

extern "C" int pp_kron_solve_f32_into(const pp_kron* k, pp_cplx a, const pp_cplxf* in, pp_cplxf* out,
                                      pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(k && in && out, "null argument");
    k->k.apply32(make_double2(a.re, a.im), reinterpret_cast<const float2*>(in), reinterpret_cast<float2*>(out),
                 (hipStream_t)stream, 1.f, nullptr, nullptr);
  });
}

extern "C" int pp_kron_solve_f32_overlap(const pp_kron* k, pp_cplx a, const pp_cplxf* in, pp_cplxf* out,
                                         pp_cplxf* sz, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(k && in && out && sz, "null argument");
    const float2* i = reinterpret_cast<const float2*>(in);
    float2* o = reinterpret_cast<float2*>(out);
    PP_REQUIRE(k->k.overlap_product32(i, o, nullptr), "the overlap product needs the one-launch form (n <= 64)");
    k->k.apply32(make_double2(a.re, a.im), i, o, (hipStream_t)stream, 1.f, nullptr, nullptr, nullptr,
                 reinterpret_cast<float2*>(sz));
  });
}

extern "C" int pp_kron_set_form(pp_kron* k, int form) {
  return guarded([&] {
    PP_REQUIRE(k && form >= 0 && form <= 2, "Kronecker form must be 0, 1 or 2");
    k->k.split32 = form >= 1;
    k->k.fused32 = form == 2;
  });
}

extern "C" int pp_cayley_set_kron(pp_cayley* c, const pp_kron* k) {
  return guarded([&] {
    PP_REQUIRE(c, "null argument");
    PP_REQUIRE(!k || (long long)k->k.nchan * k->k.n * k->k.n == c->c.n, "Kronecker shape does not match psi");
    c->c.kron = k ? &k->k : nullptr;
  });
}

extern "C" int pp_mixed_create(pp_hamiltonian* ham, const pp_kron* k, int basis_size, double tol_rel,
                               double inner_rtol, int max_outer, pp_mixed** out) {
  return guarded([&] {
    PP_REQUIRE(ham && k && out, "null argument");
    PP_REQUIRE(basis_size >= 1 && basis_size <= 60, "inner basis size must be in 1..60");
    PP_REQUIRE(tol_rel > 0 && inner_rtol > 0 && max_outer >= 1, "invalid tolerances");
    PP_REQUIRE((long long)k->k.nchan * k->k.n * k->k.n == ham->h.n, "Kronecker shape does not match psi");
    PP_REQUIRE(ham->h.fused, "the mixed solver needs the fused banded Hamiltonian");
    auto* p = new pp_mixed();
    p->c.ham = &ham->h;
    p->c.kron = &k->k;
    p->c.m = basis_size;
    p->c.tol_rel = tol_rel;
    p->c.inner_rtol = inner_rtol;
    p->c.max_outer = max_outer;
    *out = p;
  });
}

extern "C" int pp_mixed_destroy(pp_mixed* m) {
  return guarded([&] { delete m; });
}

extern "C" int pp_mixed_set_learned(pp_mixed* m, pp_learned* l) {
  return guarded([&] {
    PP_REQUIRE(m, "null argument");
    if (l) PP_REQUIRE(l->L.nchan == m->c.kron->nchan && l->L.n == m->c.kron->n, "network shape does not match");
    m->c.learned = l ? &l->L : nullptr;
  });
}

extern "C" int pp_mixed_set_fallback(pp_mixed* m, pp_cayley* plain, int fallback_iterations) {
  return guarded([&] {
    PP_REQUIRE(m && fallback_iterations >= 1, "invalid argument");
    PP_REQUIRE(!plain || plain->c.n == m->c.ham->n, "fallback solver size does not match");
    m->c.fallback = plain ? &plain->c : nullptr;
    m->c.fallback_iterations = fallback_iterations;
  });
}

extern "C" int pp_mixed_set_learned_mode(pp_mixed* m, int mode) {
  return guarded([&] {
    PP_REQUIRE(m && mode >= 0 && mode <= 2, "learned mode must be 0, 1 or 2");
    m->c.learned_mode = mode;
  });
}

extern "C" int pp_mixed_set_gram_schmidt(pp_mixed* m, int passes) {
  return guarded([&] {
    PP_REQUIRE(m && (passes == 1 || passes == 2), "Gram-Schmidt passes must be 1 or 2");
    m->c.work.gs_passes = passes;
  });
}

extern "C" int pp_mixed_set_overlap_from_solve(pp_mixed* m, int on) {
  return guarded([&] {
    PP_REQUIRE(m && (on == 0 || on == 1), "invalid argument");
    m->c.overlap_from_solve = on != 0;
  });
}

extern "C" int pp_mixed_set_deferred_finish(pp_mixed* m, int on) {
  return guarded([&] {
    PP_REQUIRE(m && (on == 0 || on == 1), "invalid argument");
    m->c.work.defer_finish = on != 0;
  });
}

extern "C" int pp_mixed_set_graph(pp_mixed* m, int graph_mode) {
  return guarded([&] {
    PP_REQUIRE(m, "null argument");
    m->c.graph_mode = graph_mode != 0;
    m->c.graph_ok = true;
  });
}

extern "C" int pp_mixed_graph_stats(const pp_mixed* m, int* captures, int* launches) {
  return guarded([&] {
    PP_REQUIRE(m && captures && launches, "null argument");
    *captures = m->c.graph_captures;
    *launches = m->c.graph_launches;
  });
}

extern "C" int pp_mixed_set_device_mode(pp_mixed* m, int device_mode) {
  return guarded([&] {
    PP_REQUIRE(m, "null argument");
    m->c.device_mode = device_mode != 0;
  });
}

extern "C" int pp_mixed_set_factored(pp_mixed* m, const pp_fham* f) {
  return guarded([&] {
    PP_REQUIRE(m, "null argument");
    PP_REQUIRE(!f || (long long)f->f.nc * f->f.n * f->f.n == m->c.ham->n, "factored operator shape does not match");
    PP_REQUIRE(!f || f->f.npot == (int)m->c.ham->terms.size(), "factored operator has a different potential count");
    m->c.fham = f ? &f->f : nullptr;
    // the FP32 inner operator forms S (x) S in Cholesky form inside the factored multiply
    // when the representation's two radial overlap factors match it
    const Repr& R = *m->c.ham->repr;
    if (f && f->f.csd < 0 && f->f.sep_enabled && R.ndim == 3 && !R.ov[0] && R.ov[1] && R.ov[2] &&
        R.ov[1]->sd == R.ov[2]->sd && R.ov[1]->n == f->f.n && R.ov[2]->n == f->f.n && R.ov[1]->sd <= 2 * f->f.bw)
      f->f.set_cholesky32(R.ov[1]->u_host, R.ov[2]->u_host, R.ov[1]->sd);
  });
}

extern "C" int pp_cayley_set_factored(pp_cayley* c, const pp_fham* f) {
  return guarded([&] {
    PP_REQUIRE(c, "null argument");
    PP_REQUIRE(!f || (long long)f->f.nc * f->f.n * f->f.n == c->c.n, "factored operator shape does not match");
    PP_REQUIRE(!f || f->f.npot == (int)c->c.ham->terms.size(), "factored operator has a different potential count");
    c->c.fham = f ? &f->f : nullptr;
  });
}

extern "C" int pp_mixed_step(pp_mixed* m, pp_cplx* psi, double dt, const double* scalings, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(m && psi && scalings, "null argument");
    m->c.step(reinterpret_cast<double2*>(psi), dt, scalings, (hipStream_t)stream);
  });
}

extern "C" int pp_mixed_stats(const pp_mixed* m, int* counts, double* residuals) {
  return guarded([&] {
    PP_REQUIRE(m, "null argument");
    const MixedStats& st = m->c.stats;
    if (counts) {
      counts[0] = st.outer_iterations;
      counts[1] = st.inner_iterations;
      counts[2] = st.outer_ops;
      counts[3] = st.inner_ops;
      counts[4] = st.precond_applies;
      counts[5] = st.learned_applies;
      counts[6] = st.fallback;
    }
    if (residuals) {
      residuals[0] = st.residual;
      residuals[1] = st.rel_residual;
    }
  });
}

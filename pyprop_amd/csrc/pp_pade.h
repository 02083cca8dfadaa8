// Scaling-and-squaring Pade exponential of A = -i dt H for a small m x m matrix
// (pamp.h:430-565), executed by one whole workgroup (blockDim.x <= 1024 threads).
// Hm, work: column-major (col * m + row); work holds 6 m*m matrices; out_col[i] = E(i, 0);
// *err = E(m-1, 0); piv_ws: m ints.  The caller synchronises the workgroup before the call
// (Hm complete); the function ends with a barrier.
#pragma once

#include <hip/hip_runtime.h>

#include "pp_common.h"

namespace pp {

__device__ inline void pade_exp_wg(const double2* __restrict__ Hm, int m, double2 dt, int order, double2* __restrict__ work,
                            double2* __restrict__ out_col, double2* __restrict__ err, int* __restrict__ piv_ws) {
  const int mm = m * m;
  double2* A = work;
  double2* A2 = work + mm;
  double2* P = work + 2 * mm;
  double2* Q = work + 3 * mm;
  double2* T = work + 4 * mm;
  double2* E = work + 5 * mm;
  __shared__ double smax[1024];
  __shared__ double coeff[32];
  __shared__ int s_scale;
  const double2 f = make_double2(dt.y, -dt.x);  // -i * dt
  double mx = 0;
  for (int i = threadIdx.x; i < mm; i += blockDim.x) {
    const double2 a = cmul(Hm[i], f);
    A[i] = a;
    mx = fmax(mx, sqrt(a.x * a.x + a.y * a.y));
  }
  smax[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) v = fmax(v, smax[i]);
    int s = (v > 0) ? (int)ceil(log2(v)) : 0;
    s_scale = s > 0 ? s : 0;
    coeff[0] = 1.0;
    for (int i = 0; i < order; ++i) coeff[i + 1] = coeff[i] * (double)(order - i) / (double)((i + 1) * (2 * order - i));
  }
  __syncthreads();
  const double div = ldexp(1.0, -s_scale);
  for (int i = threadIdx.x; i < mm; i += blockDim.x) A[i] = cscale(A[i], div);
  __syncthreads();
  // column-major (col * m + row); C = X * Y
  auto matmul = [&](const double2* X, const double2* Y, double2* C) {
    for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) {
      const int row = idx % m, col = idx / m;
      double2 acc = make_double2(0, 0);
      for (int k = 0; k < m; ++k) cfma(acc, X[k * m + row], Y[col * m + k]);
      C[idx] = acc;
    }
    __syncthreads();
  };
  matmul(A, A, A2);
  for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) {
    const bool d = (idx % m) == (idx / m);
    Q[idx] = make_double2(d ? coeff[order] : 0.0, 0);
    P[idx] = make_double2(d ? coeff[order - 1] : 0.0, 0);
  }
  __syncthreads();
  bool odd = true;
  for (int i = order - 2; i >= 0; --i) {
    double2* X = odd ? Q : P;
    matmul(X, A2, T);
    for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) {
      double2 v = T[idx];
      if ((idx % m) == (idx / m)) v.x += coeff[i];
      X[idx] = v;
    }
    __syncthreads();
    odd = !odd;
  }
  {
    double2* X = odd ? Q : P;
    matmul(X, A, T);
    for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) X[idx] = T[idx];
    __syncthreads();
  }
  // Q = Q - P
  for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) Q[idx] = csub(Q[idx], P[idx]);
  __syncthreads();
  // LU with partial pivoting (zgetrf: pivot by |re| + |im|), then solve Q X = P (zgetrs)
  for (int k = 0; k < m; ++k) {
    if (threadIdx.x == 0) {
      int p = k;
      double best = -1;
      for (int i = k; i < m; ++i) {
        const double v = fabs(Q[k * m + i].x) + fabs(Q[k * m + i].y);
        if (v > best) { best = v; p = i; }
      }
      piv_ws[k] = p;
    }
    __syncthreads();
    const int p = piv_ws[k];
    if (p != k) {
      for (int c = threadIdx.x; c < m; c += blockDim.x) {
        double2 t = Q[c * m + k]; Q[c * m + k] = Q[c * m + p]; Q[c * m + p] = t;
      }
    }
    __syncthreads();
    const double2 pv = Q[k * m + k];
    for (int i = k + 1 + threadIdx.x; i < m; i += blockDim.x) Q[k * m + i] = cdiv(Q[k * m + i], pv);
    __syncthreads();
    for (int idx = threadIdx.x; idx < (m - k - 1) * (m - k - 1); idx += blockDim.x) {
      const int i = k + 1 + idx % (m - k - 1);
      const int c = k + 1 + idx / (m - k - 1);
      Q[c * m + i] = csub(Q[c * m + i], cmul(Q[k * m + i], Q[c * m + k]));
    }
    __syncthreads();
  }
  // apply row interchanges to P, then L then U solves, one right-hand side per thread
  for (int col = threadIdx.x; col < m; col += blockDim.x) {
    double2* b = P + col * m;
    for (int k = 0; k < m; ++k) {
      const int p = piv_ws[k];
      if (p != k) { double2 t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int i = 0; i < m; ++i) {
      double2 acc = b[i];
      for (int k = 0; k < i; ++k) acc = csub(acc, cmul(Q[k * m + i], b[k]));
      b[i] = acc;
    }
    for (int i = m - 1; i >= 0; --i) {
      double2 acc = b[i];
      for (int k = i + 1; k < m; ++k) acc = csub(acc, cmul(Q[k * m + i], b[k]));
      b[i] = cdiv(acc, Q[i * m + i]);
    }
  }
  __syncthreads();
  // E = I + 2 Q^-1 P, negated for odd
  for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) {
    double2 v = cscale(P[idx], 2.0);
    if ((idx % m) == (idx / m)) v.x += 1.0;
    if (odd) { v.x = -v.x; v.y = -v.y; }
    E[idx] = v;
  }
  __syncthreads();
  for (int i = 0; i < s_scale; ++i) {
    matmul(E, E, T);
    for (int idx = threadIdx.x; idx < mm; idx += blockDim.x) E[idx] = T[idx];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < m; i += blockDim.x) out_col[i] = E[i];  // column 0
  if (threadIdx.x == 0) *err = E[m - 1];
  __syncthreads();
}

}  // namespace pp

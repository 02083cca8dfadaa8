// Learned correction to the Kronecker preconditioner (no reference counterpart).
//
// Network, shared by all channel blocks c (in = 2 n^2 reals of one n x n complex block):
//   z1 = W1 x + b1 + e_c   h1 = GELU(z1)       W1: hidden x in, e: nchan x hidden
//   z2 = W2 h1 + b2        h2 = GELU(z2)       W2: hidden x hidden
//   y  = W3 h2 + b3                             W3: in x hidden
// GELU(z) = z Phi(z) (exact, erf form).  The matrix products run on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; x, h1, h2 and the backward signals
// enter them rounded to BF16 (round to nearest even), bias and embedding sums and the
// Adam update stay in FP32 on FP32 master weights.
//
// As a preconditioner: P(v)_c = K_c^-1 (v_c + sigma |v| y(v_c / |v|)) with the norm |v| of
// the whole vector (in flexible GMRES the first direction is r / |r|): the network
// corrects the right-hand side of the Kronecker solve, where A K^-1 is close to the
// identity, so its errors are not amplified by the stiff part of A.  Training pairs (r, d)
// with A d = r come from the plain FP64 Cayley solve; P(r) = d exactly when
// y_c = (K d - r)_c / (sigma |r|), which is the target of block c; the loss is
// sum |y - t / sigma|^2 / rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"
#include "pp_learned.h"
#include "pp_mixed.h"

namespace pp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

static int grid_for_n(long long n) {
  long long b = (n + 255) / 256;
  return (int)std::max(1LL, std::min(b, 4096LL));
}

// ------------------------------------------------------------------------------------
// BF16 GEMM on MFMA: 64 x 64 tile per workgroup, four waves of 32 x 32, K step 32
// ------------------------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 32, GLD = GBK + 8;  // LDS row: 40 BF16 = 80 B

// Tile rows r0..r0+63, k0..k0+31 of X(r, k) = p[r*sr + k*sk] into S[r][k] (zero padded).
__device__ inline void load_tile(const uint16_t* __restrict__ p, long long sr, long long sk, int R, int r0, int k0,
                                 int ke, uint16_t* S, int t) {
  if (sk == 1) {
    const int r = t >> 2, kk = (t & 3) * 8;
    const int gr = r0 + r, gk = k0 + kk;
    if (gr < R && gk + 8 <= ke && (sr & 7) == 0) {
      *reinterpret_cast<uint4*>(&S[r * GLD + kk]) = *reinterpret_cast<const uint4*>(p + (long long)gr * sr + gk);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        S[r * GLD + kk + q] = (gr < R && gk + q < ke) ? p[(long long)gr * sr + gk + q] : (uint16_t)0;
    }
  } else {
    const int kk = t >> 3, rr = (t & 7) * 8;
    const int gk = k0 + kk, gr = r0 + rr;
    if (gk < ke && gr + 8 <= R && sr == 1 && (sk & 7) == 0) {
      const uint4 q4 = *reinterpret_cast<const uint4*>(p + (long long)gk * sk + gr);
      const uint16_t* v = reinterpret_cast<const uint16_t*>(&q4);
#pragma unroll
      for (int q = 0; q < 8; ++q) S[(rr + q) * GLD + kk] = v[q];
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        S[(rr + q) * GLD + kk] = (gk < ke && gr + q < R) ? p[(long long)(gr + q) * sr + (long long)gk * sk] : (uint16_t)0;
    }
  }
}

__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, int kchunk, const uint16_t* __restrict__ a,
                                                        long long sam, long long sak, const uint16_t* __restrict__ b,
                                                        long long sbk, long long sbn, float* __restrict__ c) {
  __shared__ __align__(16) uint16_t As[GBM * GLD];
  __shared__ __align__(16) uint16_t Bs[GBN * GLD];
  const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  float* cz = c + (long long)blockIdx.z * M * N;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kq = (lane >> 4) * 8;
  for (int k0 = kb; k0 < ke; k0 += GBK) {
    load_tile(a, sam, sak, M, m0, k0, ke, As, t);
    load_tile(b, sbn, sbk, N, n0, k0, ke, Bs, t);  // B viewed as (n, k)
    __syncthreads();
    bf16x8 av[2], bv[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) av[i] = *reinterpret_cast<const bf16x8*>(&As[(wm * 32 + i * 16 + (lane & 15)) * GLD + kq]);
#pragma unroll
    for (int j = 0; j < 2; ++j) bv[j] = *reinterpret_cast<const bf16x8*>(&Bs[(wn * 32 + j * 16 + (lane & 15)) * GLD + kq]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + r;
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < M && col < N) cz[(long long)row * N + col] = acc[i][j][r];
      }
}

int gemm_bf16(int M, int N, int K, const uint16_t* a, long long sam, long long sak, const uint16_t* b, long long sbk,
              long long sbn, float* c, int splits, hipStream_t s) {
  PP_REQUIRE(M > 0 && N > 0 && K > 0, "gemm: empty operand");
  PP_REQUIRE((N + GBN - 1) / GBN < (1 << 30) && (M + GBM - 1) / GBM < 65536, "gemm: grid too large");
  splits = std::max(1, splits);
  const int kchunk = (int)ceil_div(ceil_div(K, splits), GBK) * GBK;
  const int z = (int)ceil_div(K, kchunk);
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3((N + GBN - 1) / GBN, (M + GBM - 1) / GBM, z), dim3(256), 0, s, M, N, K,
                     kchunk, a, sam, sak, b, sbk, sbn, c);
  PP_HIP(hipGetLastError());
  return z;
}

// Split-K factor: enough workgroups for the chip (about 256), at most 32 splits and at
// least 128 of K per split.  Skinny products (the inference of a few dozen channel rows)
// split most.
constexpr int kOutSplitsMax = 4;  // split-K of the output layer (its partials are rows x in)
static int pick_splits(int M, int N, int K) {
  const long long tiles = ceil_div(M, GBM) * ceil_div(N, GBN);
  int sp = (int)std::max(1LL, 256 / std::max(1LL, tiles));
  sp = std::min(sp, 32);
  sp = std::min(sp, std::max(1, K / 128));
  return sp;
}

// ------------------------------------------------------------------------------------
// elementwise pieces
// ------------------------------------------------------------------------------------
// v = sum_p parts + bias[n] (+ emb[chan[m]][n]); z_out = v; act = gelu ? GELU(v) : v;
// h_out = bf16(act); y_out = act
__global__ void bias_act_kernel(int M, int N, const float* __restrict__ parts, int nparts,
                                const float* __restrict__ bias, const float* __restrict__ emb,
                                const int* __restrict__ chan, int nchan, float* __restrict__ z_out,
                                uint16_t* __restrict__ h_out, float* __restrict__ y_out, int use_gelu) {
  const long long total = (long long)M * N;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(idx / N), n = (int)(idx % N);
    float v = 0.f;
    for (int p = 0; p < nparts; ++p) v += parts[p * total + idx];
    v += bias[n];
    if (emb) {
      const int c = chan[m];
      if (c >= 0 && c < nchan) v += emb[(long long)c * N + n];
    }
    if (z_out) z_out[idx] = v;
    const float act = use_gelu ? gelu(v) : v;
    if (h_out) h_out[idx] = f2bf(act);
    if (y_out) y_out[idx] = act;
  }
}

// d = y - t*inv_sigma ; g = 2 d / M (FP32 and BF16) ; block partials of sum d^2
__global__ __launch_bounds__(256) void loss_grad_kernel(int M, int N, const float* __restrict__ y,
                                                        const float* __restrict__ t, float inv_sigma,
                                                        float* __restrict__ gf, uint16_t* __restrict__ gh,
                                                        double2* __restrict__ partial) {
  __shared__ double sh[4];
  const long long total = (long long)M * N;
  const float sc = 2.0f / (float)M;
  double acc = 0;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const float d = y[idx] - t[idx] * inv_sigma;
    acc += (double)d * d;
    const float g = sc * d;
    gf[idx] = g;
    gh[idx] = f2bf(g);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(sh[0] + sh[1] + sh[2] + sh[3], 0.0);
}

// dz = (sum_p parts) * GELU'(z): FP32 and BF16
__global__ void gelu_back_kernel(int M, int N, const float* __restrict__ parts, int nparts, const float* __restrict__ z,
                                 float* __restrict__ dzf, uint16_t* __restrict__ dzh) {
  const long long total = (long long)M * N;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    float v = 0.f;
    for (int p = 0; p < nparts; ++p) v += parts[p * total + idx];
    const float d = v * gelu_grad(z[idx]);
    dzf[idx] = d;
    dzh[idx] = f2bf(d);
  }
}

// out[n] = sum_m x[m][n]
__global__ void colsum_kernel(int M, int N, const float* __restrict__ x, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += x[(long long)m * N + n];
  out[n] = s;
}

// out[c][n] = sum over rows m of channel c of x[m][n]
__global__ void embgrad_kernel(int M, int N, int nchan, const float* __restrict__ x, const int* __restrict__ chan,
                               float* __restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nchan * N) return;
  const int c = (int)(idx / N), n = (int)(idx % N);
  float s = 0.f;
  for (int m = 0; m < M; ++m)
    if (chan[m] == c) s += x[(long long)m * N + n];
  out[idx] = s;
}

__global__ void to_bf16_kernel(long long n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

__global__ void adam_kernel(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Norms.  One workgroup per channel block: nrm[c] = |v_c|.
template <class T2>
__global__ __launch_bounds__(256) void block_norm_kernel(int nb, const T2* __restrict__ v, double* __restrict__ nrm) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const T2* vc = v + (long long)c * nb;
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += (double)vc[i].x * vc[i].x + (double)vc[i].y * vc[i].y;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) nrm[c] = std::sqrt(sh[0] + sh[1] + sh[2] + sh[3]);
}

// nrm[nchan] = sqrt(sum_c nrm[c]^2)  (the global norm, one thread)
__global__ void global_norm_kernel(int nchan, double* __restrict__ nrm) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0;
  for (int c = 0; c < nchan; ++c) s += nrm[c] * nrm[c];
  nrm[nchan] = std::sqrt(s);
}

// x[c] = bf16(v_c / |v|) with the global norm |v| = sqrt(sum_c nrm[c]^2), summed by every
// block in the same order; block 0 also stores it in nrm[nchan] (zero vector -> 0)
__global__ void block_pack_kernel(int nchan, int nb, const float2* __restrict__ v, double* __restrict__ nrm,
                                  uint16_t* __restrict__ x) {
  double g2 = 0;
  for (int c = 0; c < nchan; ++c) g2 += nrm[c] * nrm[c];
  const double g = std::sqrt(g2);
  if (blockIdx.x == 0 && threadIdx.x == 0) nrm[nchan] = g;
  const float inv = g > 0 ? (float)(1.0 / g) : 0.f;
  const long long total = (long long)nchan * nb;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    x[2 * i] = f2bf(v[i].x * inv);
    x[2 * i + 1] = f2bf(v[i].y * inv);
  }
}


// training rows: x_c = r_c / |r|, t_c = (kd_c - r_c) / |r| (FP32) with the global |r| = nrm[nchan];
// nrm_out[c] = |r_c| / |r|
__global__ void prepare_kernel(int nchan, int nb, const double2* __restrict__ r, const double2* __restrict__ kd,
                               const double* __restrict__ nrm, float* __restrict__ x, float* __restrict__ t,
                               float* __restrict__ nrm_out) {
  const double g = nrm[nchan];
  const double inv = g > 0 ? 1.0 / g : 0.0;
  const long long total = (long long)nchan * nb;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    x[2 * i] = (float)(r[i].x * inv);
    x[2 * i + 1] = (float)(r[i].y * inv);
    t[2 * i] = (float)((kd[i].x - r[i].x) * inv);
    t[2 * i + 1] = (float)((kd[i].y - r[i].y) * inv);
    if (i < nchan) nrm_out[i] = (float)(nrm[i] * inv);
  }
}

// ------------------------------------------------------------------------------------
// Learned
// ------------------------------------------------------------------------------------
void Learned::ensure(int rows) {
  PP_REQUIRE(rows > 0, "learned: empty batch");
  if (rows <= cap) return;
  cap = rows;
  const size_t ri = (size_t)rows * in, rh = (size_t)rows * hidden;
  xh.resize(ri);
  h1h.resize(rh);
  h2h.resize(rh);
  gh.resize(ri);
  dz1h.resize(rh);
  dz2h.resize(rh);
  z1.resize(rh);
  z2.resize(rh);
  y.resize(ri);
  gf.resize(ri);
  dzf.resize(rh);
  part.resize(std::max((size_t)splits_max * rh, (size_t)kOutSplitsMax * ri));
  loss_part.resize(1024);
}

void Learned::refresh_bf16(hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for_n((long long)hidden * in)), dim3(256), 0, s,
                     (long long)hidden * in, params.ptr + o_w1, w1h.ptr);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for_n((long long)hidden * hidden)), dim3(256), 0, s,
                     (long long)hidden * hidden, params.ptr + o_w2, w2h.ptr);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for_n((long long)in * hidden)), dim3(256), 0, s,
                     (long long)in * hidden, params.ptr + o_w3, w3h.ptr);
  PP_HIP(hipGetLastError());
}

void Learned::forward(int rows, const int* chan, hipStream_t s) {
  const long long rh = (long long)rows * hidden, ri = (long long)rows * in;
  int sp = gemm_bf16(rows, hidden, in, xh.ptr, in, 1, w1h.ptr, 1, in, part.ptr,
                     std::min(splits_max, pick_splits(rows, hidden, in)), s);
  hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for_n(rh)), dim3(256), 0, s, rows, hidden, part.ptr, sp,
                     params.ptr + o_b1, params.ptr + o_emb, chan, nchan, z1.ptr, h1h.ptr, (float*)nullptr, 1);
  sp = gemm_bf16(rows, hidden, hidden, h1h.ptr, hidden, 1, w2h.ptr, 1, hidden, part.ptr,
                 std::min(splits_max, pick_splits(rows, hidden, hidden)), s);
  hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for_n(rh)), dim3(256), 0, s, rows, hidden, part.ptr, sp,
                     params.ptr + o_b2, (const float*)nullptr, (const int*)nullptr, 0, z2.ptr, h2h.ptr, (float*)nullptr,
                     1);
  sp = gemm_bf16(rows, in, hidden, h2h.ptr, hidden, 1, w3h.ptr, 1, hidden, part.ptr,
                 std::min(kOutSplitsMax, pick_splits(rows, in, hidden)), s);
  hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for_n(ri)), dim3(256), 0, s, rows, in, part.ptr, sp,
                     params.ptr + o_b3, (const float*)nullptr, (const int*)nullptr, 0, (float*)nullptr,
                     (uint16_t*)nullptr, y.ptr, 0);
  PP_HIP(hipGetLastError());
}

double Learned::backward(int rows, const int* chan, const float* t, hipStream_t s, bool want_loss) {
  const long long rh = (long long)rows * hidden, ri = (long long)rows * in;
  const int lb = (int)std::min<long long>(1024, std::max<long long>(1, ceil_div(ri, 256)));
  hipLaunchKernelGGL(loss_grad_kernel, dim3(lb), dim3(256), 0, s, rows, in, y.ptr, t, 1.0f / sigma, gf.ptr, gh.ptr,
                     loss_part.ptr);
  // layer 3
  gemm_bf16(in, hidden, rows, gh.ptr, 1, in, h2h.ptr, hidden, 1, grads.ptr + o_w3, 1, s);
  hipLaunchKernelGGL(colsum_kernel, dim3((in + 255) / 256), dim3(256), 0, s, rows, in, gf.ptr, grads.ptr + o_b3);
  int sp = gemm_bf16(rows, hidden, in, gh.ptr, in, 1, w3h.ptr, hidden, 1, part.ptr,
                     std::min(splits_max, pick_splits(rows, hidden, in)), s);
  hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for_n(rh)), dim3(256), 0, s, rows, hidden, part.ptr, sp, z2.ptr,
                     dzf.ptr, dz2h.ptr);
  // layer 2
  hipLaunchKernelGGL(colsum_kernel, dim3((hidden + 255) / 256), dim3(256), 0, s, rows, hidden, dzf.ptr,
                     grads.ptr + o_b2);
  gemm_bf16(hidden, hidden, rows, dz2h.ptr, 1, hidden, h1h.ptr, hidden, 1, grads.ptr + o_w2, 1, s);
  sp = gemm_bf16(rows, hidden, hidden, dz2h.ptr, hidden, 1, w2h.ptr, hidden, 1, part.ptr,
                 std::min(splits_max, pick_splits(rows, hidden, hidden)), s);
  hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for_n(rh)), dim3(256), 0, s, rows, hidden, part.ptr, sp, z1.ptr,
                     dzf.ptr, dz1h.ptr);
  // layer 1
  hipLaunchKernelGGL(colsum_kernel, dim3((hidden + 255) / 256), dim3(256), 0, s, rows, hidden, dzf.ptr,
                     grads.ptr + o_b1);
  hipLaunchKernelGGL(embgrad_kernel, dim3((int)ceil_div((long long)nchan * hidden, 256)), dim3(256), 0, s, rows,
                     hidden, nchan, dzf.ptr, chan, grads.ptr + o_emb);
  gemm_bf16(hidden, in, rows, dz1h.ptr, 1, hidden, xh.ptr, in, 1, grads.ptr + o_w1, 1, s);
  PP_HIP(hipGetLastError());
  if (!want_loss) return 0.0;
  std::vector<double2> h(lb);
  PP_HIP(hipMemcpyAsync(h.data(), loss_part.ptr, lb * sizeof(double2), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
  double l = 0;
  for (auto& v : h) l += v.x;
  return l / rows;
}

void Learned::adam(float lr, float b1, float b2, float eps, hipStream_t s) {
  ++adam_t;
  const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for_n(p_count)), dim3(256), 0, s, p_count, params.ptr, grads.ptr,
                     adam_m.ptr, adam_v.ptr, lr, b1, b2, eps, c1, c2);
  PP_HIP(hipGetLastError());
  refresh_bf16(s);
}

// ------------------------------------------------------------------------------------
// Inference: one row per channel block (rows <= 48), weights streamed once
// ------------------------------------------------------------------------------------
// The training GEMM tiles 64 x 64 outputs and splits K; with a few dozen rows most of its
// tile is padding and each weight tile is loaded for one row block.  Inference instead
// covers all rows (three 16-row MFMA tiles) in every wave, so every weight element is read
// from HBM once and used for all rows: the first layer gives each wave a 16-column strip of
// a 256-wide K slice (the normalised input staged in LDS as BF16), the two dense layers a
// quarter of K of a workgroup's strips, with the operands loaded straight into registers.
// Rounding points are those of the training forward pass: x = bf16(v_c / |v|), h1 and h2
// in BF16, y in FP32.
constexpr int IMR = 48;        // rows covered (3 x 16)
constexpr int IKC = 256;       // K range of one first-layer workgroup
constexpr int IKMAX = 640;     // K of the dense layers (hidden width) the streaming path takes

union BfPack {  // eight BF16 values as one MFMA operand
  bf16x8 v;
  uint16_t u[8];
};

constexpr int INP = 128;  // FP64 partial sums of squares behind the inference input's norm
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// 16 bytes at byte offset off of the buffer behind r (range-checked: zeros past its end, so a
// masked-out lane passes kOff16 and its load needs no branch)
constexpr int kOff16 = 0x7ffffff0;  // beyond every buffer read this way (all under 2 GiB)
__device__ __forceinline__ u32x4 ld16(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// n2p[b] = sum of squares (FP64) over the b-th of INP equal contiguous ranges of the nf floats
// of v (nf < 2^29).  Each pass of a thread issues its four 16-byte loads before summing them in
// order; the range ends (nf, or v not 16-byte aligned) are summed value by value.
__global__ __launch_bounds__(256) void norm_part_kernel(long long nf, const float* __restrict__ v,
                                                        double* __restrict__ n2p) {
  __shared__ double sh[4];
  const int t = threadIdx.x;
  const int chunk = (int)(((nf + INP - 1) / INP + 3) / 4 * 4);
  const int lo = (int)blockIdx.x * chunk, hi = (int)min(nf, (long long)lo + chunk);
  double acc = 0;
  if ((reinterpret_cast<uintptr_t>(v) & 15) == 0) {
    const __amdgpu_buffer_rsrc_t vr = rsrc(v, nf * 4);
    const int hv = lo + max(0, hi - lo) / 4 * 4;  // end of the whole 16-byte groups
    for (int base = lo; base < hv; base += 4096) {
      u32x4 q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + (t + u * 256) * 4;
        q[u] = ld16(vr, i < hv ? i * 4 : kOff16);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 x = __builtin_bit_cast(float4, q[u]);
        acc += (double)x.x * x.x + (double)x.y * x.y + (double)x.z * x.z + (double)x.w * x.w;
      }
    }
    if (hv + t < hi) acc += (double)v[hv + t] * v[hv + t];
  } else {
    for (int i = lo + t; i < hi; i += 256) acc += (double)v[i] * v[i];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((t & 63) == 0) sh[t >> 6] = acc;
  __syncthreads();
  if (t == 0) n2p[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// part[slice][m][n] = sum_{k in slice} bf16(v[m][k] / |v|) W1[n][k]   (v: rows x K floats,
// K a multiple of 8, v 16-byte aligned, v and W1 under 2 GiB).  A wave issues its W1 operands first (they wait for
// nothing), then the norm partials and its share of v; |v| = sqrt(sum n2p) is summed by every
// wave in the same order, and workgroup (0, 0) stores it in gout[0].
__global__ __launch_bounds__(256) void infer_l1_kernel(int M, int N, int K, const float* __restrict__ v,
                                                       const double* __restrict__ n2p, double* __restrict__ gout,
                                                       const uint16_t* __restrict__ w, float* __restrict__ part) {
  __shared__ __align__(16) uint16_t As[IMR * (IKC + 8)];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int k0 = blockIdx.y * IKC;
  const int kn = min(IKC, K - k0);  // a multiple of 8
  const int col = (blockIdx.x * 4 + wave) * 16 + (lane & 15);
  constexpr int KS = IKC / 32;
  const __amdgpu_buffer_rsrc_t wr = rsrc(w, (long long)N * K * 2), vr = rsrc(v, (long long)M * K * 4);
  u32x4 b[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int kq = s * 32 + (lane >> 4) * 8;
    b[s] = ld16(wr, col < N && kq < kn ? (col * K + k0 + kq) * 2 : kOff16);
  }
  static_assert(INP == 128, "two norm partials per lane");
  const double p0 = n2p[lane], p1 = n2p[lane + 64];
  // v in groups of 8 consecutive values (two 16-byte loads), 6 groups per thread
  constexpr int GROUPS = IMR * IKC / 8 / 256;
  u32x4 vq[GROUPS][2];
#pragma unroll
  for (int q = 0; q < GROUPS; ++q) {
    const int gidx = t + q * 256, m = gidx / (IKC / 8), kk = (gidx % (IKC / 8)) * 8;
    const int off = m < M && kk < kn ? (m * K + k0 + kk) * 4 : kOff16;
    vq[q][0] = ld16(vr, off);
    vq[q][1] = ld16(vr, off == kOff16 ? kOff16 : off + 16);
  }
  __builtin_amdgcn_sched_barrier(0);  // every load above in flight before the norm is summed
  double g2 = p0 + p1;
  for (int off = 32; off > 0; off >>= 1) g2 += __shfl_down(g2, off, 64);
  const double g = std::sqrt(__shfl(g2, 0, 64));
  const float inv = g > 0 ? (float)(1.0 / g) : 0.f;
  if (blockIdx.x == 0 && blockIdx.y == 0 && t == 0) gout[0] = g;  // read by the Kronecker combination
#pragma unroll
  for (int q = 0; q < GROUPS; ++q) {
    const int gidx = t + q * 256, m = gidx / (IKC / 8), kk = (gidx % (IKC / 8)) * 8;
    const float4 lo = __builtin_bit_cast(float4, vq[q][0]), hi = __builtin_bit_cast(float4, vq[q][1]);
    BfPack pk;
    pk.u[0] = f2bf(lo.x * inv);
    pk.u[1] = f2bf(lo.y * inv);
    pk.u[2] = f2bf(lo.z * inv);
    pk.u[3] = f2bf(lo.w * inv);
    pk.u[4] = f2bf(hi.x * inv);
    pk.u[5] = f2bf(hi.y * inv);
    pk.u[6] = f2bf(hi.z * inv);
    pk.u[7] = f2bf(hi.w * inv);
    *reinterpret_cast<bf16x8*>(&As[m * (IKC + 8) + kk]) = pk.v;
  }
  __syncthreads();
  f32x4 acc[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    if (s * 32 >= kn) break;
    const int kq = s * 32 + (lane >> 4) * 8;
    const bf16x8 bv = __builtin_bit_cast(bf16x8, b[s]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(&As[(r * 16 + (lane & 15)) * (IKC + 8) + kq]);
      acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[r], 0, 0, 0);
    }
  }
  float* out = part + (long long)blockIdx.y * M * N;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = r * 16 + (lane >> 4) * 4 + q;
      if (row < M && col < N) out[(long long)row * N + col] = acc[r][q];
    }
}

// h[m][n] = bf16(GELU(sum_slices part + b[n] + emb[m][n]))  (the training bias_act order)
__global__ void infer_act1_kernel(int M, int N, int slices, const float* __restrict__ part,
                                  const float* __restrict__ bias, const float* __restrict__ emb,
                                  uint16_t* __restrict__ h) {
  const int total = M * N;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int n = idx % N;
    float v = 0.f;
#pragma unroll 8
    for (int p = 0; p < slices; ++p) v += part[(long long)p * total + idx];
    v += bias[n];
    v += emb[idx];  // row m is channel m
    h[idx] = f2bf(gelu(v));
  }
}

// out[m][n] = act(sum_k A[m][k] W[n][k] + b[n]) for rows <= 48 and K <= 128 S (a multiple of
// 8; A, W and out under 2 GiB).  A workgroup owns NS 16-column strips; its four waves split K, each over at most S
// 32-wide steps, and load their MFMA operands (A rows and W columns, 16 bytes per lane and
// step) straight from memory, all before the first MFMA.  The four partial tiles are added
// through LDS in wave order.  GELU_BF16: bf16(GELU(.)) into hout, else the FP32 value into fout.
constexpr int ISMAX = (IKMAX / 32 + 3) / 4;  // K steps of one wave at K = IKMAX
template <bool GELU_BF16, int NS, int S>
__global__ __launch_bounds__(256) void infer_layer_kernel(int M, int N, int K, const uint16_t* __restrict__ a,
                                                          const uint16_t* __restrict__ w,
                                                          const float* __restrict__ bias,
                                                          uint16_t* __restrict__ hout, float* __restrict__ fout) {
  __shared__ f32x4 red[3][3 * NS][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int s0 = wave * S, nsteps = (K + 31) / 32;
  const int n0 = blockIdx.x * 16 * NS;
  u32x4 av[S][3], bv[S][NS];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int kq = (s0 + s) * 32 + (lane >> 4) * 8;
    const bool kok = kq < K;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int row = r * 16 + (lane & 15);
      const bool ok = kok && row < M;
      const u32x4 x = *reinterpret_cast<const u32x4*>(ok ? a + (long long)row * K + kq : a);
      av[s][r] = ok ? x : u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int col = n0 + j * 16 + (lane & 15);
      const bool ok = kok && col < N;
      const u32x4 x = *reinterpret_cast<const u32x4*>(ok ? w + (long long)col * K + kq : w);
      bv[s][j] = ok ? x : u32x4{0u, 0u, 0u, 0u};
    }
  }
  f32x4 acc[3][NS];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[r][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (s0 + s >= nsteps) break;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int j = 0; j < NS; ++j)
        acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[s][r]),
                                                            __builtin_bit_cast(bf16x8, bv[s][j]), acc[r][j], 0, 0, 0);
  }
  if (wave > 0) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int j = 0; j < NS; ++j) red[wave - 1][r * NS + j][lane] = acc[r][j];
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int j = 0; j < NS; ++j) acc[r][j] += red[p][r * NS + j][lane];
  // range-checked buffer stores: a masked-out value goes to an offset past the end and is
  // dropped, so no store waits behind a branch
  constexpr int esz = GELU_BF16 ? 2 : 4;
  const __amdgpu_buffer_rsrc_t orr = rsrc(GELU_BF16 ? (const void*)hout : (const void*)fout, (long long)M * N * esz);
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int col = n0 + j * 16 + (lane & 15);
    const float bn = bias[min(col, N - 1)];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = r * 16 + (lane >> 4) * 4 + q;
        const int off = row < M && col < N ? (row * N + col) * esz : kOff16;
        const float v = acc[r][j][q] + bn;
        if (GELU_BF16) __builtin_amdgcn_raw_buffer_store_b16(f2bf(gelu(v)), orr, off, 0, 0);
        else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), orr, off, 0, 0);
      }
  }
}

template <bool GELU_BF16, int NS>
static void launch_infer_layer(int M, int N, int K, const uint16_t* a, const uint16_t* w, const float* bias,
                               uint16_t* hout, float* fout, hipStream_t s) {
  const int steps = (int)ceil_div((K + 31) / 32, 4);
  const dim3 grid((unsigned)ceil_div(N, 16 * NS));
  switch (steps) {
#define PP_INFER_LAYER(S)                                                                                        \
  case S:                                                                                                        \
    hipLaunchKernelGGL((infer_layer_kernel<GELU_BF16, NS, S>), grid, dim3(256), 0, s, M, N, K, a, w, bias, hout, \
                       fout);                                                                                    \
    break;
    PP_INFER_LAYER(1)
    PP_INFER_LAYER(2)
    PP_INFER_LAYER(3)
    PP_INFER_LAYER(4)
    PP_INFER_LAYER(5)
#undef PP_INFER_LAYER
    default: PP_REQUIRE(false, "learned inference: hidden width above the streaming limit");
  }
  static_assert(ISMAX == 5, "the dispatch covers 1 .. ISMAX steps per wave");
}

// y (rows = channels) from v: the norm partials, then the streaming kernels above
static void infer_streaming(Learned& L, const float2* in, hipStream_t s) {
  const int M = L.nchan, H = L.hidden, I = L.in;
  const int slices = (int)ceil_div(I, IKC);
  PP_REQUIRE((size_t)slices * M * H <= L.part.count, "learned: partial buffer too small");
  hipLaunchKernelGGL(norm_part_kernel, dim3(INP), dim3(256), 0, s, (long long)M * I,
                     reinterpret_cast<const float*>(in), L.n2p.ptr);
  hipLaunchKernelGGL(infer_l1_kernel, dim3((unsigned)ceil_div(H, 64), slices), dim3(256), 0, s, M, H, I,
                     reinterpret_cast<const float*>(in), L.n2p.ptr, L.nrm.ptr + M, L.w1h.ptr, L.part.ptr);
  hipLaunchKernelGGL(infer_act1_kernel, dim3((unsigned)ceil_div((long long)M * H, 256)), dim3(256), 0, s, M, H, slices,
                     L.part.ptr, L.params.ptr + L.o_b1, L.params.ptr + L.o_emb, L.h1h.ptr);
  launch_infer_layer<true, 1>(M, H, H, L.h1h.ptr, L.w2h.ptr, L.params.ptr + L.o_b2, L.h2h.ptr, nullptr, s);
  launch_infer_layer<false, 2>(M, I, H, L.h2h.ptr, L.w3h.ptr, L.params.ptr + L.o_b3, nullptr, L.y.ptr, s);
  PP_HIP(hipGetLastError());
}

void learned_apply(Learned& L, const Kron& K, double2 a, const float2* in, float2* out, hipStream_t s, float2* sz) {
  PP_REQUIRE(K.nchan == L.nchan && K.n == L.n, "learned preconditioner and Kronecker shapes differ");
  const int nb = L.n * L.n;
  L.ensure(L.nchan);
  const long long n = (long long)L.nchan * nb;
  if (L.nchan <= IMR && L.hidden <= IKMAX && L.hidden % 8 == 0 && L.in % 8 == 0 &&
      (reinterpret_cast<uintptr_t>(in) & 15) == 0 && (long long)L.in * std::max(L.nchan, L.hidden) * 4 < (1LL << 31)) {
    infer_streaming(L, in, s);
  } else {
    hipLaunchKernelGGL((block_norm_kernel<float2>), dim3(L.nchan), dim3(256), 0, s, nb, in, L.nrm.ptr);
    hipLaunchKernelGGL(block_pack_kernel, dim3(grid_for_n(n)), dim3(256), 0, s, L.nchan, nb, in, L.nrm.ptr, L.xh.ptr);
    L.forward(L.nchan, L.iota.ptr, s);
  }
  // out = K^-1 (in + sigma |v| y): the combination is formed as the Kronecker solve loads its input
  K.apply32_add(a, in, reinterpret_cast<const float2*>(L.y.ptr), L.nrm.ptr + L.nchan, L.sigma, out, s, sz);
  PP_HIP(hipGetLastError());
}

}  // namespace pp

// ------------------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------------------
using namespace pp;

extern "C" int pp_learned_create(int nchan, int n, int hidden, unsigned long long seed, pp_learned** out) {
  return guarded([&] {
    PP_REQUIRE(out && nchan > 0 && n > 0 && hidden > 0 && hidden % 8 == 0, "invalid learned-network shape");
    auto* p = new pp_learned();
    Learned& L = p->L;
    L.nchan = nchan;
    L.n = n;
    L.in = 2 * n * n;
    L.hidden = hidden;
    long long o = 0;
    L.o_w1 = o; o += (long long)hidden * L.in;
    L.o_b1 = o; o += hidden;
    L.o_emb = o; o += (long long)nchan * hidden;
    L.o_w2 = o; o += (long long)hidden * hidden;
    L.o_b2 = o; o += hidden;
    L.o_w3 = o; o += (long long)L.in * hidden;
    L.o_b3 = o; o += L.in;
    L.p_count = o;
    // deterministic initialisation: W1, W2 uniform in +-1/sqrt(fan_in); W3, biases, embedding zero
    std::vector<float> h(o, 0.f);
    unsigned long long st = seed ^ 0x9E3779B97F4A7C15ULL;
    auto uni = [&]() {
      st += 0x9E3779B97F4A7C15ULL;
      unsigned long long z = st;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      z ^= z >> 31;
      return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const double s1 = 1.0 / std::sqrt((double)L.in), s2 = 1.0 / std::sqrt((double)hidden);
    for (long long i = 0; i < (long long)hidden * L.in; ++i) h[L.o_w1 + i] = (float)(s1 * uni());
    for (long long i = 0; i < (long long)hidden * hidden; ++i) h[L.o_w2 + i] = (float)(s2 * uni());
    hipStream_t s = nullptr;
    L.params.upload(h.data(), h.size(), s);
    std::vector<float> zero(o, 0.f);
    L.grads.upload(zero.data(), zero.size(), s);
    L.adam_m.upload(zero.data(), zero.size(), s);
    L.adam_v.upload(zero.data(), zero.size(), s);
    L.w1h.resize((size_t)hidden * L.in);
    L.w2h.resize((size_t)hidden * hidden);
    L.w3h.resize((size_t)L.in * hidden);
    std::vector<int> iota(nchan);
    for (int c = 0; c < nchan; ++c) iota[c] = c;
    L.iota.upload(iota.data(), iota.size(), s);
    L.nrm.resize(nchan + 1);
    L.n2p.resize(INP);
    L.refresh_bf16(s);
    PP_HIP(hipStreamSynchronize(s));
    *out = p;
  });
}

extern "C" int pp_learned_destroy(pp_learned* l) {
  return guarded([&] { delete l; });
}

extern "C" int pp_learned_param_count(const pp_learned* l, long long* count) {
  return guarded([&] {
    PP_REQUIRE(l && count, "null argument");
    *count = l->L.p_count;
  });
}

extern "C" int pp_learned_get_params(const pp_learned* l, float* host) {
  return guarded([&] {
    PP_REQUIRE(l && host, "null argument");
    PP_HIP(hipMemcpy(host, l->L.params.ptr, l->L.p_count * sizeof(float), hipMemcpyDeviceToHost));
  });
}

extern "C" int pp_learned_set_params(pp_learned* l, const float* host) {
  return guarded([&] {
    PP_REQUIRE(l && host, "null argument");
    PP_HIP(hipMemcpy(l->L.params.ptr, host, l->L.p_count * sizeof(float), hipMemcpyHostToDevice));
    l->L.refresh_bf16(nullptr);
    PP_HIP(hipDeviceSynchronize());
  });
}

extern "C" int pp_learned_get_grads(const pp_learned* l, float* host) {
  return guarded([&] {
    PP_REQUIRE(l && host, "null argument");
    PP_HIP(hipMemcpy(host, l->L.grads.ptr, l->L.p_count * sizeof(float), hipMemcpyDeviceToHost));
  });
}

extern "C" int pp_learned_set_scale(pp_learned* l, float sigma) {
  return guarded([&] {
    PP_REQUIRE(l && sigma > 0, "invalid output scale");
    l->L.sigma = sigma;
  });
}

static void load_rows(Learned& L, int rows, const float* x, hipStream_t s) {
  L.ensure(rows);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for_n((long long)rows * L.in)), dim3(256), 0, s,
                     (long long)rows * L.in, x, L.xh.ptr);
}

extern "C" int pp_learned_forward(pp_learned* l, int rows, const float* x, const int* chan, float* y,
                                  pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(l && x && chan && y && rows > 0, "invalid argument");
    Learned& L = l->L;
    hipStream_t s = (hipStream_t)stream;
    load_rows(L, rows, x, s);
    L.forward(rows, chan, s);
    PP_HIP(hipMemcpyAsync(y, L.y.ptr, (size_t)rows * L.in * sizeof(float), hipMemcpyDeviceToDevice, s));
  });
}

extern "C" int pp_learned_gradient(pp_learned* l, int rows, const float* x, const int* chan, const float* t,
                                   double* loss, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(l && x && chan && t && rows > 0, "invalid argument");
    Learned& L = l->L;
    hipStream_t s = (hipStream_t)stream;
    load_rows(L, rows, x, s);
    L.forward(rows, chan, s);
    const double v = L.backward(rows, chan, t, s, loss != nullptr);
    if (loss) *loss = v;
  });
}

extern "C" int pp_learned_train_step(pp_learned* l, int rows, const float* x, const int* chan, const float* t,
                                     float lr, double* loss, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(l && x && chan && t && rows > 0 && lr > 0, "invalid argument");
    Learned& L = l->L;
    hipStream_t s = (hipStream_t)stream;
    load_rows(L, rows, x, s);
    L.forward(rows, chan, s);
    const double v = L.backward(rows, chan, t, s, loss != nullptr);
    L.adam(lr, 0.9f, 0.999f, 1e-8f, s);
    if (loss) *loss = v;
  });
}

extern "C" int pp_learned_prepare(pp_learned* l, const pp_kron* k, double dt, const pp_cplx* r, const pp_cplx* d,
                                  float* x_out, float* t_out, float* nrm_out, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(l && k && r && d && x_out && t_out && nrm_out, "null argument");
    Learned& L = l->L;
    PP_REQUIRE(k->k.nchan == L.nchan && k->k.n == L.n, "Kronecker and network shapes differ");
    hipStream_t s = (hipStream_t)stream;
    const long long n = (long long)L.nchan * L.n * L.n;
    if (L.tmp64.count < (size_t)n) L.tmp64.resize(n);
    PP_HIP(hipMemcpyAsync(L.tmp64.ptr, d, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    k->k.multiply64(make_double2(0.0, 0.5 * dt), L.tmp64.ptr, s);
    const double2* rr = reinterpret_cast<const double2*>(r);
    hipLaunchKernelGGL((block_norm_kernel<double2>), dim3(L.nchan), dim3(256), 0, s, L.n * L.n, rr, L.nrm.ptr);
    hipLaunchKernelGGL(global_norm_kernel, dim3(1), dim3(64), 0, s, L.nchan, L.nrm.ptr);
    hipLaunchKernelGGL(prepare_kernel, dim3(grid_for_n(n)), dim3(256), 0, s, L.nchan, L.n * L.n, rr, L.tmp64.ptr,
                       L.nrm.ptr, x_out, t_out, nrm_out);
    PP_HIP(hipGetLastError());
  });
}

extern "C" int pp_learned_apply(pp_learned* l, const pp_kron* k, double dt, const pp_cplxf* in, pp_cplxf* out,
                                pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(l && k && in && out, "null argument");
    learned_apply(l->L, k->k, make_double2(0.0, 0.5 * dt), reinterpret_cast<const float2*>(in),
                  reinterpret_cast<float2*>(out), (hipStream_t)stream);
  });
}

extern "C" int pp_gemm_bf16(int M, int N, int K, const uint16_t* a, long long sam, long long sak, const uint16_t* b,
                            long long sbk, long long sbn, float* c, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(a && b && c, "null argument");
    gemm_bf16(M, N, K, a, sam, sak, b, sbk, sbn, c, 1, (hipStream_t)stream);
  });
}

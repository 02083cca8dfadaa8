// Vector kernels for the Krylov solvers: deterministic two-stage reductions (fixed
// block partition, fixed summation order), block inner products for classical
// Gram-Schmidt, and the vector updates, all in FP64 on complex double2.
//
// Semantics: InnerProduct = sum conj(x) y (zdotc), Norm = sqrt(Re <x,x>),
// block InnerProduct = V^H r (zgemv 'C'), following core/krylov/piram/piram/functors.h:
// 185-241 and blitzblas.h:379-487.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"

namespace pp {

constexpr int kRedThreads = 256;
constexpr int kDotGroup = 8;   // inner products accumulated per pass over w

int reduction_blocks(long long n) {
  long long b = (n + 2047) / 2048;
  if (b < 1) b = 1;
  if (b > 512) b = 512;
  return (int)b;
}

void Reducer::ensure(int nblocks, int width) {
  size_t need = (size_t)nblocks * width;
  if (partial.count < need) partial.resize(need);
  if (result.count < (size_t)width) result.resize(std::max(width, 64));
  blocks = nblocks;
}

__device__ inline double2 block_sum(double2 v, double2* sh) {
  // wave reduction then across the 4 waves, fixed order
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double2 r = make_double2(0, 0);
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { r.x += sh[w].x; r.y += sh[w].y; }
  }
  return r;
}

// partial[b * width + j] for j in [j0, j0 + g): sum over this block's chunk of conj(V_j) w;
// when include_norm, partial[b * width + m] = sum |w|^2
__global__ __launch_bounds__(kRedThreads) void block_dot_kernel(long long n, int m, const double2* __restrict__ V,
                                                                 long long ld, const double2* __restrict__ w,
                                                                 double2* __restrict__ partial, int width,
                                                                 const int* cond) {
  if (cond && *cond == 0) return;
  __shared__ double2 sh[kRedThreads / 64];
  const long long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long long beg = chunk * blockIdx.x;
  const long long end = min(n, beg + chunk);
  for (int j0 = 0; j0 <= m; j0 += kDotGroup) {
    double2 acc[kDotGroup];
#pragma unroll
    for (int g = 0; g < kDotGroup; ++g) acc[g] = make_double2(0, 0);
    for (long long i = beg + threadIdx.x; i < end; i += blockDim.x) {
      const double2 wv = w[i];
#pragma unroll
      for (int g = 0; g < kDotGroup; ++g) {
        const int j = j0 + g;
        if (j < m) cfmac(acc[g], V[j * ld + i], wv);
        else if (j == m) { acc[g].x = fma(wv.x, wv.x, acc[g].x); acc[g].x = fma(wv.y, wv.y, acc[g].x); }
      }
    }
#pragma unroll
    for (int g = 0; g < kDotGroup; ++g) {
      const int j = j0 + g;
      if (j > m) break;
      double2 s = block_sum(acc[g], sh);
      if (threadIdx.x == 0) partial[(long long)blockIdx.x * width + j] = s;
    }
  }
}

// out[j] = sum_b partial[b * width + j]: one 256-thread workgroup per column j, strided loads
// then a tree (fixed order, so results are reproducible)
__global__ __launch_bounds__(256) void finish_kernel(const double2* __restrict__ partial, int blocks, int width,
                                                     int count, double2* __restrict__ out, const int* cond) {
  __shared__ double2 sh[256];
  if (cond && *cond == 0) return;
  const int j = blockIdx.x;
  if (j >= count) return;
  double2 s = make_double2(0, 0);
  for (int b = threadIdx.x; b < blocks; b += 256) {
    const double2 p = partial[(long long)b * width + j];
    s.x += p.x;
    s.y += p.y;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      sh[threadIdx.x].x += sh[threadIdx.x + off].x;
      sh[threadIdx.x].y += sh[threadIdx.x + off].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sh[0];
}

void block_dot_if(long long n, int m, const double2* V, long long ld, const double2* w, double2* out_dev,
                  Reducer& red, const int* cond, hipStream_t s) {
  const int blocks = reduction_blocks(n);
  const int width = m + 1;
  red.ensure(blocks, width);
  hipLaunchKernelGGL(block_dot_kernel, dim3(blocks), dim3(kRedThreads), 0, s, n, m, V, ld, w, red.partial.ptr, width,
                     cond);
  hipLaunchKernelGGL(finish_kernel, dim3(width), dim3(256), 0, s, red.partial.ptr, blocks, width, width, out_dev,
                     cond);
  PP_HIP(hipGetLastError());
}

void block_dot(long long n, int m, const double2* V, long long ld, const double2* w, double2* out_dev,
               Reducer& red, hipStream_t s) {
  block_dot_if(n, m, V, ld, w, out_dev, red, nullptr, s);
}

// w -= sum_j V_j h_j, then norm partials
__global__ __launch_bounds__(kRedThreads) void block_update_kernel(long long n, int m, const double2* __restrict__ V,
                                                                    long long ld, const double2* __restrict__ h,
                                                                    double2* __restrict__ w, double2* __restrict__ partial,
                                                                    const int* cond) {
  if (cond && *cond == 0) return;
  __shared__ double2 sh[kRedThreads / 64];
  __shared__ double2 hs[128];
  for (int j = threadIdx.x; j < m && j < 128; j += blockDim.x) hs[j] = h[j];
  __syncthreads();
  const long long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long long beg = chunk * blockIdx.x;
  const long long end = min(n, beg + chunk);
  double2 acc = make_double2(0, 0);
  for (long long i = beg + threadIdx.x; i < end; i += blockDim.x) {
    double2 t = make_double2(0, 0);
    for (int j = 0; j < m; ++j) cfma(t, V[j * ld + i], hs[j]);
    double2 wv = w[i];
    wv.x -= t.x;
    wv.y -= t.y;
    w[i] = wv;
    acc.x = fma(wv.x, wv.x, acc.x);
    acc.x = fma(wv.y, wv.y, acc.x);
  }
  double2 s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

void block_update_if(long long n, int m, const double2* V, long long ld, const double2* h_dev, double2* w,
                     double2* norm2_dev, Reducer& red, const int* cond, hipStream_t s) {
  PP_REQUIRE(m <= 128, "at most 128 Krylov vectors");
  const int blocks = reduction_blocks(n);
  red.ensure(blocks, 1);
  hipLaunchKernelGGL(block_update_kernel, dim3(blocks), dim3(kRedThreads), 0, s, n, m, V, ld, h_dev, w, red.partial.ptr,
                     cond);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, red.partial.ptr, blocks, 1, 1, norm2_dev, cond);
  PP_HIP(hipGetLastError());
}

void block_update(long long n, int m, const double2* V, long long ld, const double2* h_dev, double2* w,
                  double2* norm2_dev, Reducer& red, hipStream_t s) {
  block_update_if(n, m, V, ld, h_dev, w, norm2_dev, red, nullptr, s);
}

void vec_dot_into(long long n, const double2* x, const double2* y, double2* out_dev, double2* partial,
                  hipStream_t s) {
  // single inner product: one "row" of V; partial holds 2 reduction_blocks(n) values
  const int blocks = reduction_blocks(n);
  hipLaunchKernelGGL(block_dot_kernel, dim3(blocks), dim3(kRedThreads), 0, s, n, 1, x, n, y, partial, 2,
                     (const int*)nullptr);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, partial, blocks, 2, 1, out_dev, (const int*)nullptr);
  PP_HIP(hipGetLastError());
}

void vec_dot(long long n, const double2* x, const double2* y, double2* out_dev, Reducer& red, hipStream_t s) {
  red.ensure(reduction_blocks(n), 2);
  vec_dot_into(n, x, y, out_dev, red.partial.ptr, s);
}

void vec_norm2(long long n, const double2* x, double2* out_dev, Reducer& red, hipStream_t s) {
  const int blocks = reduction_blocks(n);
  red.ensure(blocks, 1);
  hipLaunchKernelGGL(block_dot_kernel, dim3(blocks), dim3(kRedThreads), 0, s, n, 0, x, n, x, red.partial.ptr, 1,
                     (const int*)nullptr);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, red.partial.ptr, blocks, 1, 1, out_dev, (const int*)nullptr);
  PP_HIP(hipGetLastError());
}

// y = x * f(scale): invert_real -> x / scale.x ; otherwise x * scale (complex)
__global__ void scale_copy_kernel(long long n, const double2* __restrict__ x, double2* __restrict__ y,
                                  const double2* __restrict__ sc, int invert_real) {
  const double2 c = *sc;
  const double2 f = invert_real ? make_double2(1.0 / c.x, 0.0) : c;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = cmul(x[i], f);
}

static int grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

void vec_scale_copy(long long n, const double2* x, double2* y, const double2* scale_dev, bool invert_real,
                    hipStream_t s) {
  hipLaunchKernelGGL(scale_copy_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, x, y, scale_dev, invert_real ? 1 : 0);
  PP_HIP(hipGetLastError());
}

__global__ void axpy_dev_kernel(long long n, const double2* __restrict__ a, int negate, const double2* __restrict__ x,
                                double2* __restrict__ y) {
  double2 c = *a;
  if (negate) { c.x = -c.x; c.y = -c.y; }
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double2 v = y[i];
    cfma(v, c, x[i]);
    y[i] = v;
  }
}

void vec_axpy_dev(long long n, const double2* alpha_dev, bool negate, const double2* x, double2* y, hipStream_t s) {
  hipLaunchKernelGGL(axpy_dev_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, alpha_dev, negate ? 1 : 0, x, y);
  PP_HIP(hipGetLastError());
}

__global__ void combine_kernel(long long n, double2 a, const double2* __restrict__ x, double2 b,
                               const double2* __restrict__ y, double2* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double2 v = make_double2(0, 0);
    if (x) cfma(v, a, x[i]);
    if (y) cfma(v, b, y[i]);
    out[i] = v;
  }
}

void vec_combine(long long n, double2 a, const double2* x, double2 b, const double2* y, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, a, x, b, y, out);
  PP_HIP(hipGetLastError());
}

// out = scale * sum_j coef_j V_j  (scale optional)
__global__ void lincomb_kernel(long long n, int m, const double2* __restrict__ V, long long ld,
                               const double2* __restrict__ coef, const double2* __restrict__ scale,
                               double2* __restrict__ out) {
  __shared__ double2 cs[128];
  for (int j = threadIdx.x; j < m; j += blockDim.x) cs[j] = coef[j];
  __syncthreads();
  const double2 sc = scale ? *scale : make_double2(1.0, 0.0);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double2 t = make_double2(0, 0);
    for (int j = 0; j < m; ++j) cfma(t, V[j * ld + i], cs[j]);
    out[i] = cmul(t, sc);
  }
}

void vec_lincomb(long long n, int m, const double2* V, long long ld, const double2* coef_dev,
                 const double2* scale_dev, double2* out, hipStream_t s) {
  PP_REQUIRE(m <= 128, "at most 128 Krylov vectors");
  hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, m, V, ld, coef_dev, scale_dev, out);
  PP_HIP(hipGetLastError());
}

void vec_copy(long long n, const double2* x, double2* y, hipStream_t s) {
  if (x != y) PP_HIP(hipMemcpyAsync(y, x, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
}

__global__ void zero_kernel(long long n, double2* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = make_double2(0, 0);
}

void vec_zero(long long n, double2* y, hipStream_t s) {
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, y);
  PP_HIP(hipGetLastError());
}

}  // namespace pp

using namespace pp;

extern "C" int pp_vec_dot(int64_t n, const pp_cplx* x, const pp_cplx* y, pp_cplx* result, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(x && y && result, "NULL argument");
    Reducer red;
    red.ensure(reduction_blocks(n), 2);
    vec_dot(n, reinterpret_cast<const double2*>(x), reinterpret_cast<const double2*>(y), red.result.ptr, red,
            (hipStream_t)stream);
    double2 h;
    PP_HIP(hipMemcpyAsync(&h, red.result.ptr, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
    PP_HIP(hipStreamSynchronize((hipStream_t)stream));
    result->re = h.x;
    result->im = h.y;
  });
}

extern "C" int pp_vec_norm(int64_t n, const pp_cplx* x, double* result, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(x && result, "NULL argument");
    Reducer red;
    red.ensure(reduction_blocks(n), 1);
    vec_norm2(n, reinterpret_cast<const double2*>(x), red.result.ptr, red, (hipStream_t)stream);
    double2 h;
    PP_HIP(hipMemcpyAsync(&h, red.result.ptr, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
    PP_HIP(hipStreamSynchronize((hipStream_t)stream));
    *result = std::sqrt(h.x);
  });
}

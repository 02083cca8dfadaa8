// Banded overlap of a B-spline rank: Cholesky factor S = U^H U (zpbtrf 'U') and the
// fibre-wise triangular multiplies / solves of core/representation/overlapmatrix.cpp:
//   multiply  x <- U^H (U x)        (ztbmv 'U','N' then 'C', :123-128)
//   solve     x <- U^-1 (U^-H x)    (ztbsv 'U','C' then 'N', :130-135)
//   sqrt      x <- U x | U^H x, x <- U^-1 x | U^-H x (:137-149)
// applied to every 1-D fibre of one rank of a C-ordered tensor (MapToRank3 loops,
// :155-229).  One thread owns one fibre; consecutive threads own consecutive
// trailing indices, so loads are coalesced whenever the rank is not the last one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"

namespace pp {

// U(i, j), j - sd <= i <= j, stored at u[j * (sd + 1) + sd + i - j]
__device__ inline double uel(const double* u, int sd, int i, int j) { return u[j * (sd + 1) + sd + i - j]; }

template <int OP, bool CONJ>
__global__ void overlap_fibre_kernel(const double* __restrict__ u, int n, int sd, long long pre, long long post,
                                     double2* __restrict__ data) {
  const long long fib = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (fib >= pre * post) return;
  const long long p = fib / post, q = fib % post;
  double2* x = data + p * n * post + q;
  const long long st = post;
  auto X = [&](int i) -> double2& { return x[i * st]; };
  // x <- U x (ascending, uses x_j for j >= i)
  auto mul_u = [&]() {
    for (int i = 0; i < n; ++i) {
      double2 acc = make_double2(0, 0);
      const int jmax = min(n - 1, i + sd);
      for (int j = i; j <= jmax; ++j) {
        const double a = uel(u, sd, i, j);
        const double2 v = X(j);
        acc.x = fma(a, v.x, acc.x);
        acc.y = fma(a, v.y, acc.y);
      }
      X(i) = acc;
    }
  };
  // x <- U^H x (descending, uses x_j for j <= i); U is real
  auto mul_uh = [&]() {
    for (int i = n - 1; i >= 0; --i) {
      double2 acc = make_double2(0, 0);
      const int jmin = max(0, i - sd);
      for (int j = jmin; j <= i; ++j) {
        const double a = uel(u, sd, j, i);
        const double2 v = X(j);
        acc.x = fma(a, v.x, acc.x);
        acc.y = fma(a, v.y, acc.y);
      }
      X(i) = acc;
    }
  };
  // x <- U^-1 x (back substitution)
  auto solve_u = [&]() {
    for (int i = n - 1; i >= 0; --i) {
      double2 acc = X(i);
      const int jmax = min(n - 1, i + sd);
      for (int j = i + 1; j <= jmax; ++j) {
        const double a = uel(u, sd, i, j);
        const double2 v = X(j);
        acc.x = fma(-a, v.x, acc.x);
        acc.y = fma(-a, v.y, acc.y);
      }
      const double d = uel(u, sd, i, i);
      X(i) = make_double2(acc.x / d, acc.y / d);
    }
  };
  // x <- U^-H x (forward substitution)
  auto solve_uh = [&]() {
    for (int i = 0; i < n; ++i) {
      double2 acc = X(i);
      const int jmin = max(0, i - sd);
      for (int j = jmin; j < i; ++j) {
        const double a = uel(u, sd, j, i);
        const double2 v = X(j);
        acc.x = fma(-a, v.x, acc.x);
        acc.y = fma(-a, v.y, acc.y);
      }
      const double d = uel(u, sd, i, i);
      X(i) = make_double2(acc.x / d, acc.y / d);
    }
  };
  if (OP == PP_OVERLAP_MULTIPLY) { mul_u(); mul_uh(); }
  else if (OP == PP_OVERLAP_SOLVE) { solve_uh(); solve_u(); }
  else if (OP == PP_OVERLAP_SQRT_MULTIPLY) { if (CONJ) mul_uh(); else mul_u(); }
  else { if (CONJ) solve_uh(); else solve_u(); }
}

// Same operations with the band width fixed at compile time: each thread keeps a sliding
// window of SD+1 fibre values in registers, so every element of the fibre is loaded and
// stored exactly once per pass and the recurrences run out of registers.  The band
// coefficients of row/column i are uniform across the wave (scalar loads).
template <int OP, bool CONJ, int SD>
__global__ __launch_bounds__(128) void overlap_window_kernel(const double* __restrict__ u, int n, long long pre,
                                                            long long post, double2* __restrict__ data) {
  const long long fib = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (fib >= pre * post) return;
  const long long p = fib / post, q = fib % post;
  double2* x = data + p * n * post + q;
  const long long st = post;
  auto U = [&](int i, int j) { return u[j * (SD + 1) + SD + i - j]; };
  auto ld = [&](int i) { return (i >= 0 && i < n) ? x[i * st] : make_double2(0, 0); };
  // x <- U x, ascending: x_i = sum_{d=0..SD} U(i, i+d) x_{i+d} (old values)
  auto mul_u = [&]() {
    double2 w[SD + 1];
#pragma unroll
    for (int d = 0; d <= SD; ++d) w[d] = ld(d);
    for (int i = 0; i < n; ++i) {
      double2 acc = make_double2(0, 0);
#pragma unroll
      for (int d = 0; d <= SD; ++d) {
        if (i + d < n) {
          const double a = U(i, i + d);
          acc.x = fma(a, w[d].x, acc.x);
          acc.y = fma(a, w[d].y, acc.y);
        }
      }
      x[i * st] = acc;
#pragma unroll
      for (int d = 0; d < SD; ++d) w[d] = w[d + 1];
      w[SD] = ld(i + SD + 1);
    }
  };
  // x <- U^H x, descending: x_i = sum_{d=0..SD} U(i-d, i) x_{i-d} (old values)
  auto mul_uh = [&]() {
    double2 w[SD + 1];
#pragma unroll
    for (int d = 0; d <= SD; ++d) w[d] = ld(n - 1 - d);
    for (int i = n - 1; i >= 0; --i) {
      double2 acc = make_double2(0, 0);
#pragma unroll
      for (int d = 0; d <= SD; ++d) {
        if (i - d >= 0) {
          const double a = U(i - d, i);
          acc.x = fma(a, w[d].x, acc.x);
          acc.y = fma(a, w[d].y, acc.y);
        }
      }
      x[i * st] = acc;
#pragma unroll
      for (int d = 0; d < SD; ++d) w[d] = w[d + 1];
      w[SD] = ld(i - SD - 1);
    }
  };
  // x <- U^-1 x, descending; window holds the SD most recently solved values x_{i+1..i+SD}
  auto solve_u = [&]() {
    double2 w[SD + 1];
#pragma unroll
    for (int d = 0; d <= SD; ++d) w[d] = make_double2(0, 0);
    for (int i = n - 1; i >= 0; --i) {
      double2 acc = x[i * st];
#pragma unroll
      for (int d = 1; d <= SD; ++d) {
        if (i + d < n) {
          const double a = U(i, i + d);
          acc.x = fma(-a, w[d - 1].x, acc.x);
          acc.y = fma(-a, w[d - 1].y, acc.y);
        }
      }
      const double dg = U(i, i);
      acc = make_double2(acc.x / dg, acc.y / dg);
      x[i * st] = acc;
#pragma unroll
      for (int d = SD; d > 0; --d) w[d] = w[d - 1];
      w[0] = acc;
    }
  };
  // x <- U^-H x, ascending; window holds x_{i-1..i-SD}
  auto solve_uh = [&]() {
    double2 w[SD + 1];
#pragma unroll
    for (int d = 0; d <= SD; ++d) w[d] = make_double2(0, 0);
    for (int i = 0; i < n; ++i) {
      double2 acc = x[i * st];
#pragma unroll
      for (int d = 1; d <= SD; ++d) {
        if (i - d >= 0) {
          const double a = U(i - d, i);
          acc.x = fma(-a, w[d - 1].x, acc.x);
          acc.y = fma(-a, w[d - 1].y, acc.y);
        }
      }
      const double dg = U(i, i);
      acc = make_double2(acc.x / dg, acc.y / dg);
      x[i * st] = acc;
#pragma unroll
      for (int d = SD; d > 0; --d) w[d] = w[d - 1];
      w[0] = acc;
    }
  };
  if (OP == PP_OVERLAP_MULTIPLY) { mul_u(); mul_uh(); }
  else if (OP == PP_OVERLAP_SOLVE) { solve_uh(); solve_u(); }
  else if (OP == PP_OVERLAP_SQRT_MULTIPLY) { if (CONJ) mul_uh(); else mul_u(); }
  else { if (CONJ) solve_uh(); else solve_u(); }
}

// Element-parallel triangular multiplies (no recurrence): one thread per tensor element,
// out-of-place.  mode 0: y = U x, mode 1: y = U^H x along the rank of extent n with
// stride `post`.  Used for the multiply ops, which need no sequential sweep.
template <int SD, int MODE>
__global__ __launch_bounds__(256) void overlap_elem_kernel(const double* __restrict__ u, int n, long long post,
                                                           long long total, const double2* __restrict__ x,
                                                           double2* __restrict__ y) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long long q = idx % post;
  const int i = (int)((idx / post) % n);
  const long long base = idx - (long long)i * post;  // element 0 of this fibre
  double2 acc = make_double2(0, 0);
#pragma unroll
  for (int d = 0; d <= SD; ++d) {
    const int j = MODE == 0 ? i + d : i - d;
    if (j >= 0 && j < n) {
      const double a = MODE == 0 ? u[j * (SD + 1) + SD - d] : u[i * (SD + 1) + SD - d];
      const double2 v = x[base + (long long)j * post];
      acc.x = fma(a, v.x, acc.x);
      acc.y = fma(a, v.y, acc.y);
    }
  }
  (void)q;
  y[idx] = acc;
}

template <int SD>
static void launch_window(int op, bool conj, const double* u, int n, long long pre, long long post, double2* data,
                          int blocks, int threads, hipStream_t s) {
#define PP_OW(OPC, CJ) \
  hipLaunchKernelGGL((overlap_window_kernel<OPC, CJ, SD>), dim3(blocks), dim3(threads), 0, s, u, n, pre, post, data)
  switch (op) {
    case PP_OVERLAP_MULTIPLY: PP_OW(PP_OVERLAP_MULTIPLY, false); break;
    case PP_OVERLAP_SOLVE: PP_OW(PP_OVERLAP_SOLVE, false); break;
    case PP_OVERLAP_SQRT_MULTIPLY:
      if (conj) PP_OW(PP_OVERLAP_SQRT_MULTIPLY, true); else PP_OW(PP_OVERLAP_SQRT_MULTIPLY, false);
      break;
    default:
      if (conj) PP_OW(PP_OVERLAP_SQRT_SOLVE, true); else PP_OW(PP_OVERLAP_SQRT_SOLVE, false);
  }
#undef PP_OW
}

void overlap_apply(const OverlapFactor& ov, int op, bool conj, int ndim, const int* extents, int rank,
                   double2* data, hipStream_t s, double2* work) {
  PP_REQUIRE(rank >= 0 && rank < ndim, "rank out of range");
  PP_REQUIRE(extents[rank] == ov.n, "overlap size does not match the rank extent");
  long long pre = 1, post = 1;
  for (int r = 0; r < rank; ++r) pre *= extents[r];
  for (int r = rank + 1; r < ndim; ++r) post *= extents[r];
  const long long fibres = pre * post;
  const int threads = 128;
  const int blocks = ceil_div(fibres, threads);
  if (blocks == 0) return;
  PP_REQUIRE(op >= 0 && op <= 3, "unknown overlap op");
  if (ov.sd >= 1 && ov.sd <= 8 && (op == PP_OVERLAP_MULTIPLY || op == PP_OVERLAP_SQRT_MULTIPLY)) {
    const long long total = pre * ov.n * post;
    if (!work && ov.work.count < (size_t)total) ov.work.resize(total);
    const int eb = ceil_div(total, 256);
    double2* w = work ? work : ov.work.ptr;
    auto pass = [&](int mode, const double2* in, double2* out) {
      switch (ov.sd * 2 + mode) {
#define PP_EL(K) \
  case 2 * K: hipLaunchKernelGGL((overlap_elem_kernel<K, 0>), dim3(eb), dim3(256), 0, s, ov.u.ptr, ov.n, post, total, in, out); break; \
  case 2 * K + 1: hipLaunchKernelGGL((overlap_elem_kernel<K, 1>), dim3(eb), dim3(256), 0, s, ov.u.ptr, ov.n, post, total, in, out); break;
        PP_EL(1) PP_EL(2) PP_EL(3) PP_EL(4) PP_EL(5) PP_EL(6) PP_EL(7) PP_EL(8)
#undef PP_EL
      }
    };
    if (op == PP_OVERLAP_MULTIPLY) {
      pass(0, data, w);     // U x
      pass(1, w, data);     // U^H (U x)
    } else {
      pass(conj ? 1 : 0, data, w);
      PP_HIP(hipMemcpyAsync(data, w, total * sizeof(double2), hipMemcpyDeviceToDevice, s));
    }
    PP_HIP(hipGetLastError());
    return;
  }
  if (ov.sd >= 1 && ov.sd <= 8) {
    switch (ov.sd) {
#define PP_SD(K) case K: launch_window<K>(op, conj, ov.u.ptr, ov.n, pre, post, data, blocks, threads, s); break;
      PP_SD(1) PP_SD(2) PP_SD(3) PP_SD(4) PP_SD(5) PP_SD(6) PP_SD(7) PP_SD(8)
#undef PP_SD
    }
    PP_HIP(hipGetLastError());
    return;
  }
#define PP_OV_LAUNCH(OPC, CJ) \
  hipLaunchKernelGGL((overlap_fibre_kernel<OPC, CJ>), dim3(blocks), dim3(threads), 0, s, ov.u.ptr, ov.n, ov.sd, pre, post, data)
  switch (op) {
    case PP_OVERLAP_MULTIPLY: PP_OV_LAUNCH(PP_OVERLAP_MULTIPLY, false); break;
    case PP_OVERLAP_SOLVE: PP_OV_LAUNCH(PP_OVERLAP_SOLVE, false); break;
    case PP_OVERLAP_SQRT_MULTIPLY:
      if (conj) PP_OV_LAUNCH(PP_OVERLAP_SQRT_MULTIPLY, true); else PP_OV_LAUNCH(PP_OVERLAP_SQRT_MULTIPLY, false);
      break;
    case PP_OVERLAP_SQRT_SOLVE:
      if (conj) PP_OV_LAUNCH(PP_OVERLAP_SQRT_SOLVE, true); else PP_OV_LAUNCH(PP_OVERLAP_SQRT_SOLVE, false);
      break;
    default: throw Error(PP_ERR_INVALID, "unknown overlap op");
  }
#undef PP_OV_LAUNCH
  PP_HIP(hipGetLastError());
}

// Upper banded Cholesky S = U^T U of a real symmetric positive definite band matrix
// (the arithmetic of zpbtrf('U') for a real matrix).
void band_cholesky_upper_host(int n, int sd, const double* s, std::vector<double>& u) {
  const int ld = sd + 1;
  u.assign(s, s + (size_t)n * ld);
  auto U = [&](int i, int j) -> double& { return u[(size_t)j * ld + sd + i - j]; };
  for (int j = 0; j < n; ++j) {
    const int i0 = std::max(0, j - sd);
    for (int i = i0; i <= j; ++i) {
      double sum = U(i, j);
      const int k0 = std::max(i0, std::max(0, i - sd));
      for (int k = k0; k < i; ++k) sum -= U(k, i) * U(k, j);
      if (i < j) {
        U(i, j) = sum / U(i, i);
      } else {
        if (!(sum > 0.0)) throw Error(PP_ERR_INVALID, "overlap matrix is not positive definite");
        U(j, j) = std::sqrt(sum);
      }
    }
  }
}

}  // namespace pp

using namespace pp;

extern "C" int pp_overlap_create(int n, int superdiagonals, const double* s_upper, pp_overlap** out) {
  return guarded([&] {
    PP_REQUIRE(out && s_upper && n > 0 && superdiagonals >= 0 && superdiagonals < n, "invalid overlap arguments");
    auto* o = new pp_overlap();
    try {
      o->f.n = n;
      o->f.sd = superdiagonals;
      band_cholesky_upper_host(n, superdiagonals, s_upper, o->f.u_host);
      o->f.u.upload(o->f.u_host.data(), o->f.u_host.size());
      PP_HIP(hipDeviceSynchronize());
    } catch (...) {
      delete o;
      throw;
    }
    *out = o;
  });
}

extern "C" int pp_overlap_get_factor(const pp_overlap* ov, double* u_upper) {
  return guarded([&] {
    PP_REQUIRE(ov && u_upper, "NULL argument");
    std::copy(ov->f.u_host.begin(), ov->f.u_host.end(), u_upper);
  });
}

extern "C" int pp_overlap_destroy(pp_overlap* ov) {
  return guarded([&] { delete ov; });
}

extern "C" int pp_overlap_apply(const pp_overlap* ov, int op, int conj, int ndim, const int* extents, int rank,
                                pp_cplx* data, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(ov && extents && data, "NULL argument");
    overlap_apply(ov->f, op, conj != 0, ndim, extents, rank, reinterpret_cast<double2*>(data), (hipStream_t)stream);
  });
}

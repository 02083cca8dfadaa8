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

void overlap_apply(const OverlapFactor& ov, int op, bool conj, int ndim, const int* extents, int rank,
                   double2* data, hipStream_t s) {
  PP_REQUIRE(rank >= 0 && rank < ndim, "rank out of range");
  PP_REQUIRE(extents[rank] == ov.n, "overlap size does not match the rank extent");
  long long pre = 1, post = 1;
  for (int r = 0; r < rank; ++r) pre *= extents[r];
  for (int r = rank + 1; r < ndim; ++r) post *= extents[r];
  const long long fibres = pre * post;
  const int threads = 128;
  const int blocks = ceil_div(fibres, threads);
  if (blocks == 0) return;
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
static void band_cholesky_upper(int n, int sd, const double* s, std::vector<double>& u) {
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
      band_cholesky_upper(n, superdiagonals, s_upper, o->f.u_host);
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

// Potential generation on the device: grid functions represented in B-spline pairs.
//
// GeneratePotential (pyprop/tensorpotential/TensorPotentialGenerator.py:80-274) evaluates
// a potential on the quadrature grid of every B-spline rank, then represents it in the
// basis one rank at a time, last rank first (:165-230).  Each rank is
// RepresentPotentialInBasisBSpline (core/tensorpotential/basis_bspline.h:10-47): with the
// source viewed as (pre, Q, post) (MapToRank3), every (pre, pair, post) entry is
// BSplineGlobalOverlapIntegral (core/transform/bspline/bspline.cpp:372-466) of the grid
// vector src(pre, :, post), scaled by 1/2 on the diagonal pairs of "Herm" storage.
//
// Design.  A projection table holds, per pair p, the first grid point lo[p] of the common
// support of its two splines, the support length len[p] <= k*g, and the weights
//   W[s][p] = w_q B_i(q) B_j^(d)(q) * (-1)^d [row > col] * (1/2 [Herm diagonal])
// for q = lo[p] + s, (i, j) = (min, max)(row, col), built once on the host from the
// spline tables.  One rank is then the banded reduction
//   dst[a][p][c] = sum_{s < len[p]} W[s][p] src[a][lo[p] + s][c],
// at most k*g = 49 terms per entry (k = 7, g = 7).  A workgroup stages 8 grid vectors
// in LDS (8 consecutive rows when post = 1, 8 consecutive columns otherwise) and a thread
// computes one pair for all 8, so each source element is read from HBM once and each
// weight load feeds 8 FMAs.  Measured on helium R12 (394 angular pairs): the last rank
// with the grid evaluated in place, 0.83 ms (964 MB written); the first rank, 3.1 ms
// (964 MB read, 3.6 GB complex written).  Sixteen vectors per workgroup, a transposed
// [q][v] tile read 16 bytes at a time and chunked weight loads were all slower (fewer
// resident waves).
//
// Grid functions on the device:
//  * multipole: F[a][q1][q2] = sum_l C[l][a] r<^l / r>^(l+1), the 1/r12 expansion that
//    CustomPotentialEvaluationR12 evaluates on the (q1, q2) grid
//    (examples/tensor/helium_stabilization/coupledspherical.cpp:60-111);
//  * separable: F[a][q1][q2] = sum_t C[t][a] f1[t][q1] f2[t][q2], the one-electron
//    potentials (kinetic, centrifugal, Coulomb, length-gauge dipole) of potential.cpp and
//    coupledlength.cpp, whose grid values the caller passes per rank.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"

namespace pp {

// A workgroup projects kProjTile grid vectors, staged in LDS as [vector][Q + 1].
constexpr int kProjTile = 8;
constexpr int kProjMaxQ = 2000;  // 8 x 2001 doubles of LDS (125 KB)

struct BsProjection {
  int Q = 0, P = 0, L = 0;
  DeviceBuffer<int> lo, len;
  DeviceBuffer<double> W;  // [s][p]
};

__device__ inline void proj_store(double* dst, long long i, double v, int accumulate) {
  dst[i] = accumulate ? dst[i] + v : v;
}
__device__ inline void proj_store(double2* dst, long long i, double v, int accumulate) {
  if (accumulate) {
    double2 o = dst[i];
    o.x += v;
    dst[i] = o;
  } else {
    dst[i] = make_double2(v, 0.0);
  }
}

// acc[v] = sum_{s < n} W[s * P] tile[v][l0 + s] for the vectors v of the tile (one weight
// load feeds kProjTile FMAs).
__device__ inline void tile_dot(int l0, int n, const double* __restrict__ W, int P, const double* tile, int ld,
                                double (&acc)[kProjTile]) {
#pragma unroll
  for (int v = 0; v < kProjTile; ++v) acc[v] = 0.0;
  for (int s = 0; s < n; ++s) {
    const double w = W[(long long)s * P];
    const double* t = tile + l0 + s;
#pragma unroll
    for (int v = 0; v < kProjTile; ++v) acc[v] = fma(w, t[v * ld], acc[v]);
  }
}

// Sources of the rows kernel: row r of the (rows, Q) grid function.
struct ArraySource {
  const double* src;
  int Q;
  __device__ double operator()(long long r, int q) const { return src[r * Q + q]; }
};
// row r = a * n1 + q1 of F[a][q1][q2] = sum_l C[l][a] r<^lam_l / r>^(lam_l + 1)
struct MultipoleSource {
  long long P0;
  int nlam, n1;
  const int* lams;
  const double* C;
  const double* r1;
  const double* r2;
  __device__ double operator()(long long r, int q2) const {
    const long long a = r / n1;
    const int q1 = (int)(r - a * n1);
    const double x1 = r1[q1], x2 = r2[q2];
    const double rlo = fmin(x1, x2), rhi = fmax(x1, x2);
    const double ratio = rlo / rhi;
    double pw = 1.0, v = 0.0;
    int cur = 0;
    for (int l = 0; l < nlam; ++l) {
      for (; cur < lams[l]; ++cur) pw *= ratio;
      v = fma(C[(long long)l * P0 + a], pw, v);
    }
    return v / rhi;
  }
};
// row r = a * n1 + q1 of F[a][q1][q2] = sum_t C[t][a] f1[t][q1] f2[t][q2] (f1 = NULL:
// F[a][q2] = sum_t C[t][a] f2[t][q2], one radial rank)
struct SeparableSource {
  long long P0;
  int terms, n1, n2;
  const double* C;
  const double* f1;
  const double* f2;
  __device__ double operator()(long long r, int q2) const {
    const long long a = r / n1;
    const int q1 = (int)(r - a * n1);
    double v = 0.0;
    for (int t = 0; t < terms; ++t) {
      const double g1 = f1 ? f1[(long long)t * n1 + q1] : 1.0;  // no f1: one radial rank
      v = fma(C[(long long)t * P0 + a] * g1, f2[(long long)t * n2 + q2], v);
    }
    return v;
  }
};

// post == 1: the grid axis is contiguous.  Block = kProjTile consecutive rows; thread = pair.
// The rows come from memory or are evaluated in place (the grid of GeneratePotential's
// step 4 is then never stored).
template <typename OUT, typename SRC>
__global__ __launch_bounds__(256) void project_rows_kernel(long long pre, int Q, int P, const int* __restrict__ lo,
                                                           const int* __restrict__ len,
                                                           const double* __restrict__ W, SRC src,
                                                           OUT* __restrict__ dst, int accumulate) {
  extern __shared__ double tile[];  // kProjTile x (Q + 1)
  const int ld = Q + 1;
  const long long a0 = (long long)blockIdx.x * kProjTile;
  const int rows = (int)min((long long)kProjTile, pre - a0);
  for (int r = 0; r < kProjTile; ++r)
    for (int q = threadIdx.x; q < Q; q += blockDim.x) tile[r * ld + q] = r < rows ? src(a0 + r, q) : 0.0;
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    double acc[kProjTile];
    tile_dot(lo[p], len[p], W + p, P, tile, ld, acc);
    for (int r = 0; r < rows; ++r) proj_store(dst, (a0 + r) * P + p, acc[r], accumulate);
  }
}

// post > 1: the grid axis has stride post.  Block = one a and kProjTile consecutive
// columns c; thread = pair, holding the columns' sums.
template <typename OUT>
__global__ __launch_bounds__(256) void project_cols_kernel(long long pre, int Q, long long post, int P,
                                                           const int* __restrict__ lo, const int* __restrict__ len,
                                                           const double* __restrict__ W,
                                                           const double* __restrict__ src, OUT* __restrict__ dst,
                                                           int accumulate) {
  extern __shared__ double tile[];  // kProjTile x (Q + 1)
  const int ld = Q + 1;
  const long long ntiles = (post + kProjTile - 1) / kProjTile;
  const long long a = blockIdx.x / ntiles;
  const long long c0 = (blockIdx.x - a * ntiles) * kProjTile;
  const int cols = (int)min((long long)kProjTile, post - c0);
  for (int idx = threadIdx.x; idx < kProjTile * Q; idx += blockDim.x) {
    const int q = idx / kProjTile, c = idx - q * kProjTile;
    tile[c * ld + q] = c < cols ? src[(a * Q + q) * post + c0 + c] : 0.0;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    double acc[kProjTile];
    tile_dot(lo[p], len[p], W + p, P, tile, ld, acc);
    OUT* out = dst + (a * P + p) * post + c0;
    if (cols == kProjTile) {
#pragma unroll
      for (int c = 0; c < kProjTile; ++c) proj_store(out, c, acc[c], accumulate);
    } else {
      for (int c = 0; c < cols; ++c) proj_store(out, c, acc[c], accumulate);
    }
  }
}

// F[a][q1][q2] = sum_l C[l][a] r<^lam_l / r>^(lam_l + 1), lams ascending.
__global__ void multipole_grid_kernel(long long P0, int nlam, const int* __restrict__ lams,
                                      const double* __restrict__ C, int n1, const double* __restrict__ r1, int n2,
                                      const double* __restrict__ r2, double* __restrict__ F) {
  const long long total = P0 * n1 * n2;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long a = idx / ((long long)n1 * n2);
    const int rem = (int)(idx - a * n1 * n2);
    const int q1 = rem / n2, q2 = rem - q1 * n2;
    const double x1 = r1[q1], x2 = r2[q2];
    const double rlo = fmin(x1, x2), rhi = fmax(x1, x2);
    const double ratio = rlo / rhi;
    double pw = 1.0, v = 0.0;
    int cur = 0;
    for (int l = 0; l < nlam; ++l) {
      for (; cur < lams[l]; ++cur) pw *= ratio;
      v = fma(C[(long long)l * P0 + a], pw, v);
    }
    F[idx] = v / rhi;
  }
}

// F[a][q1][q2] = sum_t C[t][a] f1[t][q1] f2[t][q2]
__global__ void separable_grid_kernel(long long P0, int terms, const double* __restrict__ C, int n1,
                                      const double* __restrict__ f1, int n2, const double* __restrict__ f2,
                                      double* __restrict__ F) {
  const long long total = P0 * n1 * n2;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long a = idx / ((long long)n1 * n2);
    const int rem = (int)(idx - a * n1 * n2);
    const int q1 = rem / n2, q2 = rem - q1 * n2;
    double v = 0.0;
    for (int t = 0; t < terms; ++t) v = fma(C[(long long)t * P0 + a] * f1[(long long)t * n1 + q1], f2[(long long)t * n2 + q2], v);
    F[idx] = v;
  }
}

static int grid_blocks(long long total) { return (int)std::max<long long>(1, std::min<long long>(65536, (total + 255) / 256)); }

}  // namespace pp

struct pp_bspline_projection {
  pp::BsProjection p;
};

using namespace pp;

extern "C" int pp_bspline_projection_create(int n_splines, int n_points, int order, const double* weights,
                                            const double* values, const double* deriv_values,
                                            const int* support_first, const int* support_count,
                                            int differentiation, int pair_count, const int* pairs, int hermitian,
                                            pp_bspline_projection** out) {
  return guarded([&] {
    PP_REQUIRE(out && weights && values && deriv_values && support_first && support_count && pairs,
               "null argument");
    PP_REQUIRE(n_splines > 0 && n_points > 0 && order > 0 && pair_count > 0, "invalid sizes");
    PP_REQUIRE(n_points <= kProjMaxQ, "too many grid points for the projection tile");
    PP_REQUIRE(differentiation >= 0 && differentiation <= 2, "differentiation must be 0, 1 or 2");
    const int N = n_splines, Q = n_points, P = pair_count;
    for (int i = 0; i < N; ++i)
      PP_REQUIRE(support_first[i] >= 0 && support_count[i] >= 0 && support_first[i] + support_count[i] <= Q,
                 "spline support outside the grid");
    std::vector<int> lo(P, 0), len(P, 0);
    int L = 1;
    for (int p = 0; p < P; ++p) {
      int row = pairs[2 * p], col = pairs[2 * p + 1];
      PP_REQUIRE(row >= 0 && row < N && col >= 0 && col < N, "pair index outside the basis");
      const int i = std::min(row, col), j = std::max(row, col);
      if (j - i >= order) continue;
      const int a = std::max(support_first[i], support_first[j]);
      const int b = std::min(support_first[i] + support_count[i], support_first[j] + support_count[j]);
      lo[p] = a;
      len[p] = std::max(0, b - a);
      L = std::max(L, len[p]);
    }
    std::vector<double> W((size_t)L * P, 0.0);
    for (int p = 0; p < P; ++p) {
      const int row = pairs[2 * p], col = pairs[2 * p + 1];
      const int i = std::min(row, col), j = std::max(row, col);
      double factor = (row > col && (differentiation & 1)) ? -1.0 : 1.0;
      if (hermitian && row == col) factor *= 0.5;
      for (int s = 0; s < len[p]; ++s) {
        const int q = lo[p] + s;
        W[(size_t)s * P + p] = factor * weights[q] * values[(size_t)i * Q + q] * deriv_values[(size_t)j * Q + q];
      }
    }
    auto* h = new pp_bspline_projection();
    h->p.Q = Q;
    h->p.P = P;
    h->p.L = L;
    h->p.lo.upload(lo.data(), lo.size());
    h->p.len.upload(len.data(), len.size());
    h->p.W.upload(W.data(), W.size());
    PP_HIP(hipDeviceSynchronize());  // the uploads read host vectors that go out of scope
    *out = h;
  });
}

extern "C" int pp_bspline_projection_destroy(pp_bspline_projection* p) {
  return guarded([&] { delete p; });
}

extern "C" int pp_bspline_projection_info(const pp_bspline_projection* p, int* n_points, int* pair_count,
                                          int* max_support) {
  return guarded([&] {
    PP_REQUIRE(p, "null argument");
    if (n_points) *n_points = p->p.Q;
    if (pair_count) *pair_count = p->p.P;
    if (max_support) *max_support = p->p.L;
  });
}

namespace pp {

template <typename OUT, typename SRC>
static void launch_rows(const BsProjection& P, long long pre, SRC src, OUT* dst, int accumulate, hipStream_t s) {
  const size_t lds = (size_t)kProjTile * (P.Q + 1) * sizeof(double);
  const long long blocks = (pre + kProjTile - 1) / kProjTile;
  PP_REQUIRE(blocks <= 0x7fffffffLL, "too many rows");
  PP_HIP(hipFuncSetAttribute((const void*)project_rows_kernel<OUT, SRC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds));
  hipLaunchKernelGGL((project_rows_kernel<OUT, SRC>), dim3((unsigned)blocks), dim3(256), lds, s, pre, P.Q, P.P,
                     P.lo.ptr, P.len.ptr, P.W.ptr, src, dst, accumulate);
  PP_HIP(hipGetLastError());
}

template <typename OUT>
static void launch_cols(const BsProjection& P, long long pre, long long post, const double* src, OUT* dst,
                        int accumulate, hipStream_t s) {
  const size_t lds = (size_t)kProjTile * (P.Q + 1) * sizeof(double);
  const long long blocks = pre * ((post + kProjTile - 1) / kProjTile);
  PP_REQUIRE(blocks <= 0x7fffffffLL, "too many column tiles");
  PP_HIP(hipFuncSetAttribute((const void*)project_cols_kernel<OUT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds));
  hipLaunchKernelGGL(project_cols_kernel<OUT>, dim3((unsigned)blocks), dim3(256), lds, s, pre, P.Q, post, P.P,
                     P.lo.ptr, P.len.ptr, P.W.ptr, src, dst, accumulate);
  PP_HIP(hipGetLastError());
}

}  // namespace pp

extern "C" int pp_bspline_represent(const pp_bspline_projection* proj, int64_t pre, int64_t post, const double* src,
                                    void* dst, int dst_complex, int accumulate, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(proj && src && dst, "null argument");
    PP_REQUIRE(pre > 0 && post > 0, "invalid sizes");
    const BsProjection& P = proj->p;
    const hipStream_t s = (hipStream_t)stream;
    if (post == 1) {
      const ArraySource in{src, P.Q};
      if (dst_complex)
        launch_rows(P, pre, in, reinterpret_cast<double2*>(dst), accumulate, s);
      else
        launch_rows(P, pre, in, reinterpret_cast<double*>(dst), accumulate, s);
    } else {
      if (dst_complex)
        launch_cols(P, pre, post, src, reinterpret_cast<double2*>(dst), accumulate, s);
      else
        launch_cols(P, pre, post, src, reinterpret_cast<double*>(dst), accumulate, s);
    }
  });
}

extern "C" int pp_potential_represent_multipole(const pp_bspline_projection* proj1,
                                                const pp_bspline_projection* proj2, int64_t pair_count, int nlam,
                                                const int* lams, const double* coeffs, const double* r1,
                                                const double* r2, double* work, pp_cplx* dst, int accumulate,
                                                pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(proj1 && proj2 && lams && coeffs && r1 && r2 && work && dst, "null argument");
    PP_REQUIRE(pair_count > 0 && nlam > 0, "invalid sizes");
    const BsProjection& A = proj1->p;
    const BsProjection& B = proj2->p;
    const hipStream_t s = (hipStream_t)stream;
    const MultipoleSource in{(long long)pair_count, nlam, A.Q, lams, coeffs, r1, r2};
    launch_rows(B, (long long)pair_count * A.Q, in, work, 0, s);
    launch_cols(A, (long long)pair_count, (long long)B.P, work, reinterpret_cast<double2*>(dst), accumulate, s);
  });
}

extern "C" int pp_potential_represent_separable(const pp_bspline_projection* proj1,
                                                const pp_bspline_projection* proj2, int64_t pair_count, int terms,
                                                const double* coeffs, const double* f1, const double* f2,
                                                double* work, pp_cplx* dst, int accumulate, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(proj1 && coeffs && f1 && dst, "null argument");
    PP_REQUIRE(pair_count > 0 && terms > 0, "invalid sizes");
    const BsProjection& A = proj1->p;
    const hipStream_t s = (hipStream_t)stream;
    if (!proj2) {
      // one radial rank: row a of F[a][q] = sum_t C[t][a] f1[t][q]
      const SeparableSource in{(long long)pair_count, terms, 1, A.Q, coeffs, nullptr, f1};
      launch_rows(A, (long long)pair_count, in, reinterpret_cast<double2*>(dst), accumulate, s);
      return;
    }
    PP_REQUIRE(f2 && work, "null argument");
    const BsProjection& B = proj2->p;
    const SeparableSource in{(long long)pair_count, terms, A.Q, B.Q, coeffs, f1, f2};
    launch_rows(B, (long long)pair_count * A.Q, in, work, 0, s);
    launch_cols(A, (long long)pair_count, (long long)B.P, work, reinterpret_cast<double2*>(dst), accumulate, s);
  });
}

extern "C" int pp_potential_grid_multipole(int64_t pair_count, int nlam, const int* lams, const double* coeffs,
                                           int n1, const double* r1, int n2, const double* r2, double* grid,
                                           pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(lams && coeffs && r1 && r2 && grid, "null argument");
    PP_REQUIRE(pair_count > 0 && nlam > 0 && n1 > 0 && n2 > 0, "invalid sizes");
    const long long total = (long long)pair_count * n1 * n2;
    hipLaunchKernelGGL(multipole_grid_kernel, dim3(grid_blocks(total)), dim3(256), 0, (hipStream_t)stream,
                       (long long)pair_count, nlam, lams, coeffs, n1, r1, n2, r2, grid);
    PP_HIP(hipGetLastError());
  });
}

extern "C" int pp_potential_grid_separable(int64_t pair_count, int terms, const double* coeffs, int n1,
                                           const double* f1, int n2, const double* f2, double* grid,
                                           pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(coeffs && f1 && f2 && grid, "null argument");
    PP_REQUIRE(pair_count > 0 && terms > 0 && n1 > 0 && n2 > 0, "invalid sizes");
    const long long total = (long long)pair_count * n1 * n2;
    hipLaunchKernelGGL(separable_grid_kernel, dim3(grid_blocks(total)), dim3(256), 0, (hipStream_t)stream,
                       (long long)pair_count, terms, coeffs, n1, f1, n2, f2, grid);
    PP_HIP(hipGetLastError());
  });
}

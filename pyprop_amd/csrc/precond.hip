// Block-diagonal ("radial") preconditioner of the Cayley step for rank-3 (two-electron)
// wavefunctions: for every angular channel c an exact solve of
//   M_c = scaling_s (S1 x S2) + scaling_h P_c        (n1*n2 x n1*n2)
// where P_c is the angular-diagonal part of H (RadialTwoElectronPreconditioner.Setup,
// examples/tensor/helium_stabilization/preconditioner.py:113-146; the SuperLU variant,
// :156-198, solves the same systems exactly).
//
// M_c is block banded: n2 x n2 blocks (I, J) with |I - J| <= b (b = B-spline order - 1).
// The host factors it once per channel by a right-looking block LU without pivoting
// (the Hermitian part of M_c is positive definite: S x S is, and scaling_h P_c is
// i dt/2 times a real symmetric matrix), and stores for each block row I
//   L_II^-1,  L'_IJ = L_II^-1 L_IJ (J = I-b..I-1),  U_II^-1,  U'_IJ = U_II^-1 U_IJ.
// The device solve is then
//   c_I = L_II^-1 b_I        (all I in parallel)
//   y_I = c_I - sum_J L'_IJ y_J   (sequential over I, one workgroup per channel, LDS resident)
//   d_I = U_II^-1 y_I        (all I in parallel)
//   x_I = d_I - sum_J U'_IJ x_J   (sequential, descending I)
#include <hip/hip_runtime.h>

#include <complex>
#include <thread>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"

namespace pp {

using cd = std::complex<double>;

namespace {

// dense n x n row-major helpers
void lu_nopivot(int n, cd* a) {
  for (int k = 0; k < n; ++k) {
    const cd piv = a[k * n + k];
    if (piv == cd(0)) throw Error(PP_ERR_INVALID, "zero pivot in preconditioner block LU");
    for (int i = k + 1; i < n; ++i) {
      cd l = a[i * n + k] / piv;
      a[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
    }
  }
}

// X U = A, U upper (from a packed LU), in place on A
void right_upper_solve(int n, const cd* lu, cd* a) {
  for (int r = 0; r < n; ++r) {
    cd* row = a + (size_t)r * n;
    for (int j = 0; j < n; ++j) {
      cd s = row[j];
      for (int k = 0; k < j; ++k) s -= row[k] * lu[k * n + j];
      row[j] = s / lu[j * n + j];
    }
  }
}

// L X = A, L unit lower (from a packed LU), in place on A
void left_unit_lower_solve(int n, const cd* lu, cd* a) {
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < i; ++k) {
      const cd l = lu[i * n + k];
      if (l == cd(0)) continue;
      for (int j = 0; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
    }
}

// C -= A B
void gemm_sub(int n, const cd* A, const cd* B, cd* C) {
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const cd a = A[i * n + k];
      if (a == cd(0)) continue;
      const cd* b = B + (size_t)k * n;
      cd* c = C + (size_t)i * n;
      for (int j = 0; j < n; ++j) c[j] -= a * b[j];
    }
}

// C = A B
void gemm(int n, const cd* A, const cd* B, cd* C) {
  std::fill(C, C + (size_t)n * n, cd(0));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const cd a = A[i * n + k];
      if (a == cd(0)) continue;
      const cd* b = B + (size_t)k * n;
      cd* c = C + (size_t)i * n;
      for (int j = 0; j < n; ++j) c[j] += a * b[j];
    }
}

void unit_lower_inverse(int n, const cd* lu, cd* out) {
  std::fill(out, out + (size_t)n * n, cd(0));
  for (int i = 0; i < n; ++i) out[i * n + i] = 1;
  left_unit_lower_solve(n, lu, out);
}

void upper_inverse(int n, const cd* lu, cd* out) {
  // solve U X = I by back substitution, column by column
  std::fill(out, out + (size_t)n * n, cd(0));
  for (int c = 0; c < n; ++c) {
    for (int i = n - 1; i >= 0; --i) {
      cd s = (i == c) ? cd(1) : cd(0);
      for (int k = i + 1; k < n; ++k) s -= lu[i * n + k] * out[k * n + c];
      out[i * n + c] = s / lu[i * n + i];
    }
  }
}

struct HostFactors {
  std::vector<cd> linv, lp, uinv, up;
};

}  // namespace

static void factor_channel(int c, int n1, int n2, int b, const double* s1, const double* s2, const cd* pdata,
                           int p1, const std::vector<int>& pos1, int p2, const std::vector<int>& pos2, cd sh,
                           cd ss, HostFactors& F) {
  const size_t bs = (size_t)n2 * n2;
  const int nb = 2 * b + 1;
  std::vector<cd> A((size_t)n1 * nb * bs, cd(0));
  auto blk = [&](int I, int J) -> cd* { return A.data() + ((size_t)I * nb + (J - I + b)) * bs; };
  const cd* pc = pdata + (size_t)c * p1 * p2;
  for (int I = 0; I < n1; ++I)
    for (int J = std::max(0, I - b); J <= std::min(n1 - 1, I + b); ++J) {
      cd* a = blk(I, J);
      const double sIJ = s1[(size_t)I * n1 + J];
      const int q1 = pos1[(size_t)I * n1 + J];
      for (int i = 0; i < n2; ++i)
        for (int j = std::max(0, i - b); j <= std::min(n2 - 1, i + b); ++j) {
          cd v = ss * (sIJ * s2[(size_t)i * n2 + j]);
          const int q2 = pos2[(size_t)i * n2 + j];
          if (q1 >= 0 && q2 >= 0) v += sh * pc[(size_t)q1 * p2 + q2];
          a[i * n2 + j] = v;
        }
    }
  // right-looking block LU
  for (int K = 0; K < n1; ++K) {
    cd* akk = blk(K, K);
    lu_nopivot(n2, akk);
    const int hi = std::min(n1 - 1, K + b);
    for (int I = K + 1; I <= hi; ++I) right_upper_solve(n2, akk, blk(I, K));
    for (int J = K + 1; J <= hi; ++J) left_unit_lower_solve(n2, akk, blk(K, J));
    for (int I = K + 1; I <= hi; ++I)
      for (int J = K + 1; J <= hi; ++J) gemm_sub(n2, blk(I, K), blk(K, J), blk(I, J));
  }
  // derived blocks
  std::vector<cd> tmp(bs);
  for (int I = 0; I < n1; ++I) {
    cd* li = F.linv.data() + ((size_t)c * n1 + I) * bs;
    cd* ui = F.uinv.data() + ((size_t)c * n1 + I) * bs;
    unit_lower_inverse(n2, blk(I, I), li);
    upper_inverse(n2, blk(I, I), ui);
    for (int q = 0; q < b; ++q) {
      const int J = I - b + q;
      cd* dst = F.lp.data() + ((size_t)c * n1 + I) * b * bs;  // [row][q*n2 + j]
      if (J >= 0) gemm(n2, li, blk(I, J), tmp.data());
      else std::fill(tmp.begin(), tmp.end(), cd(0));
      for (int i = 0; i < n2; ++i)
        for (int j = 0; j < n2; ++j) dst[(size_t)i * b * n2 + q * n2 + j] = tmp[i * n2 + j];
    }
    for (int q = 0; q < b; ++q) {
      const int J = I + 1 + q;
      cd* dst = F.up.data() + ((size_t)c * n1 + I) * b * bs;
      if (J < n1) gemm(n2, ui, blk(I, J), tmp.data());
      else std::fill(tmp.begin(), tmp.end(), cd(0));
      for (int i = 0; i < n2; ++i)
        for (int j = 0; j < n2; ++j) dst[(size_t)i * b * n2 + q * n2 + j] = tmp[i * n2 + j];
    }
  }
}

// out_I = B_I x_I for all blocks I of all channels; 8 threads per row
__global__ __launch_bounds__(512) void pc_diag_kernel(const double2* __restrict__ B, double2* __restrict__ psi, int n2) {
  extern __shared__ double2 xs[];
  const long long blkid = blockIdx.x;  // c * n1 + I
  double2* x = psi + blkid * n2;
  for (int j = threadIdx.x; j < n2; j += blockDim.x) xs[j] = x[j];
  __syncthreads();
  const double2* M = B + blkid * (long long)n2 * n2;
  const int part = threadIdx.x & 7;
  for (int row = threadIdx.x >> 3; row < n2; row += blockDim.x >> 3) {
    double2 acc = make_double2(0, 0);
    for (int j = part; j < n2; j += 8) cfma(acc, M[(long long)row * n2 + j], xs[j]);
    for (int off = 4; off > 0; off >>= 1) {
      acc.x += __shfl_xor(acc.x, off, 8);
      acc.y += __shfl_xor(acc.y, off, 8);
    }
    if (part == 0) x[row] = acc;
  }
}

// Sequential block sweep over one channel held in LDS (generic fallback, any n2).
__global__ __launch_bounds__(512) void pc_sweep_kernel(const double2* __restrict__ Mp, double2* __restrict__ psi, int n1,
                                                       int n2, int b, int forward) {
  extern __shared__ double2 y[];
  const int c = blockIdx.x;
  double2* g = psi + (long long)c * n1 * n2;
  for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) y[i] = g[i];
  __syncthreads();
  const int part = threadIdx.x & 7;
  const int width = b * n2;
  for (int step = 0; step < n1; ++step) {
    const int I = forward ? step : n1 - 1 - step;
    const double2* M = Mp + ((long long)c * n1 + I) * (long long)b * n2 * n2;
    for (int row = threadIdx.x >> 3; row < n2; row += blockDim.x >> 3) {
      double2 acc = make_double2(0, 0);
      const double2* mrow = M + (long long)row * width;
      for (int q = 0; q < b; ++q) {
        const int J = forward ? I - b + q : I + 1 + q;
        if (J < 0 || J >= n1) continue;
        const double2* yj = y + J * n2;
        for (int j = part; j < n2; j += 8) cfma(acc, mrow[q * n2 + j], yj[j]);
      }
      for (int off = 4; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, 8);
        acc.y += __shfl_xor(acc.y, off, 8);
      }
      if (part == 0) {
        double2 v = y[I * n2 + row];
        y[I * n2 + row] = make_double2(v.x - acc.x, v.y - acc.y);
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) g[i] = y[i];
}

// Block sweep for n2 <= 64: 16 threads per row, 1024 threads; all B*4 matrix elements a
// thread needs for a block step are loaded before any is used, so the step costs one
// memory latency plus the transfer of the block row instead of a chain of dependent
// loads.  The matrix is streamed with non-temporal loads (read once per solve).
typedef double pc_d2v __attribute__((ext_vector_type(2)));
template <int B>
__global__ __launch_bounds__(1024) void pc_sweep16_kernel(const double2* __restrict__ Mp, double2* __restrict__ psi,
                                                          int n1, int n2, int forward) {
  constexpr int JPER = 4;  // ceil(64 / 16)
  extern __shared__ double2 y[];
  const int c = blockIdx.x;
  double2* g = psi + (long long)c * n1 * n2;
  for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) y[i] = g[i];
  __syncthreads();
  const int part = threadIdx.x & 15;
  const int row = threadIdx.x >> 4;
  const bool active = row < n2;
  const int width = B * n2;
  for (int step = 0; step < n1; ++step) {
    const int I = forward ? step : n1 - 1 - step;
    const double2* mrow = Mp + ((long long)c * n1 + I) * (long long)B * n2 * n2 + (long long)row * width;
    double2 m[B][JPER];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int J = forward ? I - B + q : I + 1 + q;
#pragma unroll
      for (int u = 0; u < JPER; ++u) {
        const int j = part + 16 * u;
        if (active && J >= 0 && J < n1 && j < n2) {
          const pc_d2v v = __builtin_nontemporal_load(reinterpret_cast<const pc_d2v*>(mrow + q * n2 + j));
          m[q][u] = make_double2(v.x, v.y);
        } else {
          m[q][u] = make_double2(0, 0);
        }
      }
    }
    double2 acc = make_double2(0, 0);
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int J = forward ? I - B + q : I + 1 + q;
      if (J < 0 || J >= n1) continue;
      const double2* yj = y + J * n2;
#pragma unroll
      for (int u = 0; u < JPER; ++u) {
        const int j = part + 16 * u;
        if (j < n2) cfma(acc, m[q][u], yj[j]);
      }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
      acc.x += __shfl_xor(acc.x, off, 16);
      acc.y += __shfl_xor(acc.y, off, 16);
    }
    if (active && part == 0) {
      const double2 v = y[I * n2 + row];
      y[I * n2 + row] = make_double2(v.x - acc.x, v.y - acc.y);
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) g[i] = y[i];
}

static void launch_sweep(const Precond& pc, const double2* M, double2* psi, int forward, hipStream_t s) {
  const size_t lds = (size_t)pc.n1 * pc.n2 * sizeof(double2);
  if (pc.n2 <= 64 && pc.b >= 1 && pc.b <= 8) {
    switch (pc.b) {
#define PP_SW(BB) case BB: hipLaunchKernelGGL(pc_sweep16_kernel<BB>, dim3(pc.nchan), dim3(1024), lds, s, M, psi, pc.n1, pc.n2, forward); break;
      PP_SW(1) PP_SW(2) PP_SW(3) PP_SW(4) PP_SW(5) PP_SW(6) PP_SW(7) PP_SW(8)
#undef PP_SW
    }
  } else {
    hipLaunchKernelGGL(pc_sweep_kernel, dim3(pc.nchan), dim3(512), lds, s, M, psi, pc.n1, pc.n2, pc.b, forward);
  }
}

void precond_solve(const Precond& pc, double2* psi, hipStream_t s) {
  const int nblk = pc.nchan * pc.n1;
  const size_t lds_diag = pc.n2 * sizeof(double2);
  const size_t lds_sweep = (size_t)pc.n1 * pc.n2 * sizeof(double2);
  (void)lds_sweep;
  hipLaunchKernelGGL(pc_diag_kernel, dim3(nblk), dim3(512), lds_diag, s, pc.linv.ptr, psi, pc.n2);
  launch_sweep(pc, pc.lp.ptr, psi, 1, s);
  hipLaunchKernelGGL(pc_diag_kernel, dim3(nblk), dim3(512), lds_diag, s, pc.uinv.ptr, psi, pc.n2);
  launch_sweep(pc, pc.up.ptr, psi, 0, s);
  PP_HIP(hipGetLastError());
}

}  // namespace pp

using namespace pp;

extern "C" int pp_precond_create(int nchan, int n1, int n2, int band, const double* s1, const double* s2,
                                 const pp_cplx* p_data, int p1_count, const int32_t* pairs1, int p2_count,
                                 const int32_t* pairs2, pp_cplx scaling_h, pp_cplx scaling_s, pp_precond** out) {
  return guarded([&] {
    PP_REQUIRE(out && s1 && s2 && p_data && pairs1 && pairs2, "NULL argument");
    PP_REQUIRE(nchan > 0 && n1 > 0 && n2 > 0 && band >= 0 && band < n1, "invalid sizes");
    PP_REQUIRE((size_t)n1 * n2 * sizeof(double2) <= 64 * 1024, "channel block exceeds 64 KiB of LDS");
    std::vector<int> pos1((size_t)n1 * n1, -1), pos2((size_t)n2 * n2, -1);
    for (int p = 0; p < p1_count; ++p) {
      const int r = pairs1[2 * p], c = pairs1[2 * p + 1];
      PP_REQUIRE(r >= 0 && r < n1 && c >= 0 && c < n1, "pair out of range");
      pos1[(size_t)r * n1 + c] = p;
    }
    for (int p = 0; p < p2_count; ++p) {
      const int r = pairs2[2 * p], c = pairs2[2 * p + 1];
      PP_REQUIRE(r >= 0 && r < n2 && c >= 0 && c < n2, "pair out of range");
      pos2[(size_t)r * n2 + c] = p;
    }
    const size_t bs = (size_t)n2 * n2;
    HostFactors F;
    F.linv.assign((size_t)nchan * n1 * bs, 0);
    F.uinv.assign((size_t)nchan * n1 * bs, 0);
    F.lp.assign((size_t)nchan * n1 * band * bs, 0);
    F.up.assign((size_t)nchan * n1 * band * bs, 0);
    const cd sh(scaling_h.re, scaling_h.im), ss(scaling_s.re, scaling_s.im);
    const cd* pd = reinterpret_cast<const cd*>(p_data);
    unsigned nthreads = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> pool;
    std::vector<std::string> errors(nthreads);
    for (unsigned t = 0; t < nthreads; ++t) {
      pool.emplace_back([&, t] {
        try {
          for (int c = (int)t; c < nchan; c += (int)nthreads)
            factor_channel(c, n1, n2, band, s1, s2, pd, p1_count, pos1, p2_count, pos2, sh, ss, F);
        } catch (const std::exception& e) {
          errors[t] = e.what();
        }
      });
    }
    for (auto& th : pool) th.join();
    for (auto& e : errors)
      if (!e.empty()) throw Error(PP_ERR_INVALID, e);
    auto* p = new pp_precond();
    try {
      p->p.nchan = nchan;
      p->p.n1 = n1;
      p->p.n2 = n2;
      p->p.b = band;
      static_assert(sizeof(cd) == sizeof(double2), "complex layout");
      p->p.linv.upload(reinterpret_cast<const double2*>(F.linv.data()), F.linv.size());
      p->p.uinv.upload(reinterpret_cast<const double2*>(F.uinv.data()), F.uinv.size());
      p->p.lp.upload(reinterpret_cast<const double2*>(F.lp.data()), F.lp.size());
      p->p.up.upload(reinterpret_cast<const double2*>(F.up.data()), F.up.size());
      PP_HIP(hipDeviceSynchronize());
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

extern "C" int pp_precond_destroy(pp_precond* pc) {
  return guarded([&] { delete pc; });
}

extern "C" int pp_precond_solve(const pp_precond* pc, pp_cplx* psi, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(pc && psi, "NULL argument");
    precond_solve(pc->p, reinterpret_cast<double2*>(psi), (hipStream_t)stream);
  });
}

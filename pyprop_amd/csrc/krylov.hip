// Combined representation, Hamiltonian (sum of tensor potentials), GMRES, pAMP and the
// Cayley step.
//
// Reference algorithms restated here:
//   CombinedRepresentation overlap ops / inner product  core/representation/combinedrepresentation.cpp:165-238,423-474
//   BasisPropagator multiply variants                  pyprop/tensorpotential/BasisPropagator.py:116-159
//   GMRES::SolveVector / PerformArnoldiStep / DGKS      core/krylov/gmres/gmres.h:260-505
//   pAMP::PropagateVector / ScalingPadeExponential      core/krylov/pamp/pamp/pamp.h:270-670
//   CayleyPropagator.AdvanceStep / MultiplySpH          pyprop/propagator/CayleyPropagator.py:71-150
//
// GMRES runs its Arnoldi vectors and all vector work on the device; after each Arnoldi
// step the (j+2) Gram-Schmidt coefficients and two norms come back to the host, which
// does the (k+1) x k least-squares problem (Householder QR, as zgels does) and the
// convergence test, as in the reference.  pAMP keeps everything on the device,
// including the scaling-and-squaring Pade exponential of the Hessenberg matrix (one
// workgroup), so a step issues no host synchronisation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"
#include "pp_pade.h"

namespace pp {

using cd = std::complex<double>;

// ----------------------------------------------------------------------------------
// representation
// ----------------------------------------------------------------------------------
void Repr::overlap(int op, bool conj, double2* psi, hipStream_t s, double2* work) const {
  for (int r = 0; r < ndim; ++r)
    if (ov[r]) overlap_apply(*ov[r], op, conj, ndim, extents.data(), r, psi, s, work);
}

// ----------------------------------------------------------------------------------
// Hamiltonian
// ----------------------------------------------------------------------------------
void Hamiltonian::apply(const double* scalings, const double2* src, double2* dst, const double2* addend,
                        double2 alpha, hipStream_t s) {
  const int nt = (int)terms.size();
  if (fused) {
    std::vector<double2> sc(nt);
    for (int i = 0; i < nt; ++i) sc[i] = make_double2(scalings ? scalings[i] : 1.0, 0.0);
    tp_run_f64(fused_plan, terms, sc.data(), src, dst, addend, alpha, false, s);
    return;
  }
  vec_zero(n, dst, s);
  for (int i = 0; i < nt; ++i) {
    double2 sc = make_double2(scalings ? scalings[i] : 1.0, 0.0);
    tp_run_f64(terms[i]->plan, {terms[i]}, &sc, src, dst, nullptr, make_double2(1.0, 0.0), true, s);
  }
  if (addend || alpha.x != 1.0 || alpha.y != 0.0) vec_combine(n, alpha, dst, make_double2(1.0, 0.0), addend, dst, s);
}

void Hamiltonian::apply_f32(const double* scalings, const float2* src, float2* dst, const float2* addend,
                            float2 alpha, hipStream_t s) {
  PP_REQUIRE(fused, "FP32 Hamiltonian needs fusable banded potentials");
  const int nt = (int)terms.size();
  std::vector<float2> sc(nt);
  for (int i = 0; i < nt; ++i) sc[i] = make_float2((float)(scalings ? scalings[i] : 1.0), 0.f);
  tp_run_f32(fused_plan, terms, sc.data(), src, dst, addend, alpha, false, s);
}

void Hamiltonian::multiply_balanced(const double* scalings, const double2* src, double2* dst, hipStream_t s) {
  ensure_work();
  vec_copy(n, src, work.ptr, s);
  repr->overlap(PP_OVERLAP_SQRT_SOLVE, false, work.ptr, s);
  apply(scalings, work.ptr, dst, nullptr, make_double2(1.0, 0.0), s);
  repr->overlap(PP_OVERLAP_SQRT_SOLVE, true, dst, s);
}

void Hamiltonian::ensure_work() {
  if (work.count < (size_t)n) work.resize(n);
}

// ----------------------------------------------------------------------------------
// small dense helpers (host)
// ----------------------------------------------------------------------------------
// Least squares min ||A x - b|| for A (rows x cols), rows >= cols, full rank, by
// Householder QR (the method of zgels).  A column-major.
static void least_squares(int rows, int cols, std::vector<cd> A, std::vector<cd> b, std::vector<cd>& x) {
  for (int k = 0; k < cols; ++k) {
    double nrm = 0;
    for (int i = k; i < rows; ++i) nrm += std::norm(A[k * rows + i]);
    nrm = std::sqrt(nrm);
    if (nrm == 0) continue;
    cd akk = A[k * rows + k];
    cd alpha = (std::abs(akk) == 0) ? cd(-nrm, 0) : -akk / std::abs(akk) * nrm;
    std::vector<cd> v(rows - k);
    v[0] = akk - alpha;
    for (int i = k + 1; i < rows; ++i) v[i - k] = A[k * rows + i];
    double vn = 0;
    for (auto& z : v) vn += std::norm(z);
    if (vn == 0) continue;
    for (int c = k; c < cols; ++c) {
      cd d = 0;
      for (int i = k; i < rows; ++i) d += std::conj(v[i - k]) * A[c * rows + i];
      d *= 2.0 / vn;
      for (int i = k; i < rows; ++i) A[c * rows + i] -= d * v[i - k];
    }
    cd d = 0;
    for (int i = k; i < rows; ++i) d += std::conj(v[i - k]) * b[i];
    d *= 2.0 / vn;
    for (int i = k; i < rows; ++i) b[i] -= d * v[i - k];
  }
  x.assign(cols, 0);
  for (int k = cols - 1; k >= 0; --k) {
    cd sum = b[k];
    for (int c = k + 1; c < cols; ++c) sum -= A[c * rows + k] * x[c];
    x[k] = sum / A[k * rows + k];
  }
}

// ----------------------------------------------------------------------------------
// GMRES
// ----------------------------------------------------------------------------------
Gmres::Gmres(long long n_, int m_, double tol_, bool dorth_) : n(n_), m(m_), tol(tol_), dorth(dorth_) {
  PP_REQUIRE(m >= 2 && m <= 127, "GMRES basis size must be 2..127");
  if (tol < 1.0e-13) tol = 1.0e-13;  // gmres.h:224-229
  V.resize((size_t)m * n);
  r.resize(n);
  small.resize(2 * (m + 4));
  y.assign(m + 1, 0);
  err_list.assign(m, 0);
}

void Gmres::read_small(int count, hipStream_t s) {
  host_small.resize(count);
  PP_HIP(hipMemcpyAsync(host_small.data(), small.ptr, count * sizeof(double2), hipMemcpyDeviceToHost, s));
  PP_HIP(hipStreamSynchronize(s));
}

void Gmres::solve(pp_operator_fn op, void* ctx, const double2* rhs, double2* x, hipStream_t s) {
  auto apply_op = [&](const double2* in, double2* out) {
    ++op_count;
    int rc = op(ctx, reinterpret_cast<const pp_cplx*>(in), reinterpret_cast<pp_cplx*>(out), (pp_stream)s);
    if (rc != 0) throw Error(PP_ERR_INTERNAL, "GMRES operator callback failed");
  };
  op_count = 0;
  orth_count = 0;
  std::fill(err_list.begin(), err_list.end(), 0.0);
  // Residual = rhs; inputNorm = ||Residual||; Residual *= 1/inputNorm
  vec_copy(n, rhs, r.ptr, s);
  vec_norm2(n, r.ptr, small.ptr, red, s);
  read_small(1, s);
  const double input_norm = std::sqrt(host_small[0].x);
  vec_combine(n, make_double2(1.0 / input_norm, 0), r.ptr, make_double2(0, 0), nullptr, r.ptr, s);
  // PerformInitialArnoldiStep (gmres.h:260-294)
  H.assign((size_t)m * (m + 1), 0);
  auto Hc = [&](int col, int row) -> cd& { return H[(size_t)col * (m + 1) + row]; };
  bool happy = false;
  int step = 0;
  double2* V0 = V.ptr;
  vec_copy(n, r.ptr, V0, s);
  apply_op(V0, r.ptr);
  vec_dot(n, V0, r.ptr, small.ptr, red, s);
  vec_axpy_dev(n, small.ptr, true, V0, r.ptr, s);
  vec_norm2(n, r.ptr, small.ptr + 1, red, s);
  read_small(2, s);
  const cd alpha(host_small[0].x, host_small[0].y);
  Hc(0, 0) = alpha;
  double beta = std::sqrt(host_small[1].x);
  if (std::abs(beta) < tol) happy = true;
  Hc(0, 1) = beta;
  if (happy) y[0] = input_norm / alpha;
  double err = std::abs(Hc(step, step + 1));
  err_list[step] = err;
  while (step < m - 1 && !happy && err > tol) {
    // PerformArnoldiStep (gmres.h:296-354)
    const int j = step;
    double2* vj1 = V.ptr + (size_t)(j + 1) * n;
    const cd nrm = Hc(j, j + 1);
    vec_combine(n, make_double2(1.0 / nrm.real(), 0), r.ptr, make_double2(0, 0), nullptr, vj1, s);
    apply_op(vj1, r.ptr);
    const int k = j + 2;  // vectors 0..j+1
    block_dot(n, k, V.ptr, n, r.ptr, small.ptr, red, s);          // small[0..k-1] = V^H r, small[k] = |r|^2
    block_update(n, k, V.ptr, n, small.ptr, r.ptr, small.ptr + (k + 1), red, s);
    read_small(k + 2, s);
    std::vector<cd> h(k);
    for (int i = 0; i < k; ++i) h[i] = cd(host_small[i].x, host_small[i].y);
    double orig = std::sqrt(host_small[k].x);
    double resid = std::sqrt(host_small[k + 1].x);
    if (dorth) {
      int cur = 0;
      while (resid < 0.717 * orig) {
        ++orth_count;
        ++cur;
        if (cur > max_orth) break;
        block_dot(n, k, V.ptr, n, r.ptr, small.ptr, red, s);
        block_update(n, k, V.ptr, n, small.ptr, r.ptr, small.ptr + (k + 1), red, s);
        read_small(k + 2, s);
        for (int i = 0; i < k; ++i) h[i] += cd(host_small[i].x, host_small[i].y);
        orig = resid;
        resid = std::sqrt(host_small[k + 1].x);
      }
    }
    for (int i = 0; i < k; ++i) Hc(j + 1, i) = h[i];
    beta = resid;
    if (std::abs(beta) < tol) happy = true;
    Hc(j + 1, j + 2) = beta;
    ++step;
    // minimise residual (gmres.h:467-490)
    const int ks = step + 1;
    std::vector<cd> A((size_t)(ks + 1) * ks), b(ks + 1, 0), xs;
    for (int c = 0; c < ks; ++c)
      for (int rr = 0; rr <= ks; ++rr) A[(size_t)c * (ks + 1) + rr] = Hc(c, rr);
    b[0] = input_norm;
    least_squares(ks + 1, ks, A, b, xs);
    for (int i = 0; i < ks; ++i) y[i] = xs[i];
    std::vector<cd> res(ks + 1, 0);
    for (int rr = 0; rr <= ks; ++rr) {
      cd acc = 0;
      for (int c = 0; c < ks; ++c) acc += Hc(c, rr) * xs[c];
      res[rr] = acc;
    }
    res[0] -= input_norm;
    double e2 = 0;
    for (auto& z : res) e2 += std::norm(z);
    err = std::sqrt(e2);
    err_list[step] = err;
  }
  error = err;
  steps = step;
  // x = Q y over vectors 0..step (gmres.h:496-502)
  std::vector<double2> yc(step + 1);
  for (int i = 0; i <= step; ++i) yc[i] = make_double2(y[i].real(), y[i].imag());
  coef.upload(yc.data(), yc.size(), s);
  vec_lincomb(n, step + 1, V.ptr, n, coef.ptr, nullptr, x, s);
}

// ----------------------------------------------------------------------------------
// pAMP
// ----------------------------------------------------------------------------------
// scalars layout (double2): [0] |v|^2 -> norm, [1] 1/beta, [2] error estimate
// flags: [0] happy breakdown, [1] re-orthogonalisation still needed
__global__ void pamp_init_kernel(double2* sc, double2* H, int m, int* flags) {
  for (int i = threadIdx.x; i < m * m; i += blockDim.x) H[i] = make_double2(0, 0);
  if (threadIdx.x == 0) {
    const double nrm = sqrt(sc[0].x);
    sc[0] = make_double2(nrm, 0);
    sc[1] = make_double2(1.0 / nrm, 0);
    flags[0] = 0;
    flags[1] = 0;
  }
}

// PerformArnoldiStep prologue (pamp.h:305-318): beta = |r|; breakdown if beta < tol;
// H(j, j+1) = beta; scale = 1/beta
__global__ void pamp_beta_kernel(const double2* norm2, double2* sc, double2* H, int m, int j, double tol, int* flags) {
  if (flags[0]) return;
  const double beta = sqrt(norm2->x);
  if (fabs(beta) < tol) {
    flags[0] = 1;
    return;
  }
  H[j * m + (j + 1)] = make_double2(beta, 0);
  sc[1] = make_double2(1.0 / beta, 0);
}

// v_{j+1} = r * sc[1] unless broken down
__global__ void pamp_newvec_kernel(long long n, const double2* __restrict__ r, double2* __restrict__ v,
                                   const double2* sc, const int* flags) {
  if (flags[0]) return;
  const double f = sc[1].x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    v[i] = make_double2(r[i].x * f, r[i].y * f);
}

// H(j+1, 0..j+1) = h unless broken down (mode 0: flag = breakdown flag);
// mode 2: H(j+1, .) += h when the re-orthogonalisation flag is set
__global__ void pamp_store_h_kernel(const double2* h, double2* H, int m, int j, int k, const int* flag, int mode) {
  if (mode == 0 ? flag[0] != 0 : flag[0] == 0) return;
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    double2 v = h[i];
    if (mode == 2) { v.x += H[(j + 1) * m + i].x; v.y += H[(j + 1) * m + i].y; }
    H[(j + 1) * m + i] = v;
  }
}

__global__ void pamp_orth_check_kernel(const double2* orig2, const double2* resid2, int* flags) {
  flags[1] = (!flags[0] && sqrt(resid2->x) < 0.717 * sqrt(orig2->x)) ? 1 : 0;
}

// scaling-and-squaring Pade exponential of A = -i dt H (pamp.h:430-565), one workgroup.
// work: 6 m*m matrices; out_col[i] = E(i, 0); sc[2] = E(m-1, 0).
__global__ __launch_bounds__(256) void pade_exp_kernel(const double2* __restrict__ Hm, int m, double2 dt, int order,
                                                        double2* __restrict__ work, double2* __restrict__ out_col,
                                                        double2* __restrict__ sc, int* __restrict__ piv_ws) {
  pade_exp_wg(Hm, m, dt, order, work, out_col, sc + 2, piv_ws);
}

// EigenvectorExponential (pamp.h:567-608): exp(A) ~ V diag(exp(w)) V^H with the unit-norm
// eigenvectors V of A = -i dt H_m, the reference's zgeev on the m x m upper Hessenberg matrix.
// Here: complex Schur form by the single-shift QR algorithm (Givens bulge chase, Wilkinson
// shift, an exceptional shift every 11th iteration), eigenvectors of the triangular factor by
// back substitution (divisors below eps |T| raised to it, as ztrevc does), v = Q y normalised.
// h: H_m column-major (H(i, j) at [j * m + i]); returns column 0 of the exponential.
static std::vector<std::complex<double>> eigenvector_exp_col0(int m, const std::vector<double2>& h, double2 dt) {
  using cd = std::complex<double>;
  const cd f = -cd(0, 1) * cd(dt.x, dt.y);
  std::vector<cd> a((size_t)m * m), q((size_t)m * m, 0.0);  // row-major A and Q
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) a[(size_t)i * m + j] = f * cd(h[(size_t)j * m + i].x, h[(size_t)j * m + i].y);
  for (int i = 0; i < m; ++i) q[(size_t)i * m + i] = 1.0;
  auto A = [&](int i, int j) -> cd& { return a[(size_t)i * m + j]; };
  const double eps = std::numeric_limits<double>::epsilon();
  double amax = 0;
  for (const cd& z : a) amax = std::max(amax, std::abs(z));
  // apply G = [c s; -conj(s) c] to rows k, k+1 (columns >= c0), G^H to columns k, k+1 (rows <= r1) and to Q
  auto rot = [&](int k, double c, cd sn, int c0, int r1) {
    for (int j = c0; j < m; ++j) {
      const cd x = A(k, j), y = A(k + 1, j);
      A(k, j) = c * x + sn * y;
      A(k + 1, j) = c * y - std::conj(sn) * x;
    }
    for (int i = 0; i <= r1; ++i) {
      const cd x = A(i, k), y = A(i, k + 1);
      A(i, k) = c * x + std::conj(sn) * y;
      A(i, k + 1) = c * y - sn * x;
    }
    for (int i = 0; i < m; ++i) {
      const cd x = q[(size_t)i * m + k], y = q[(size_t)i * m + k + 1];
      q[(size_t)i * m + k] = c * x + std::conj(sn) * y;
      q[(size_t)i * m + k + 1] = c * y - sn * x;
    }
  };
  auto givens = [](cd x, cd y, double& c, cd& sn) {
    const double ax = std::abs(x), ay = std::abs(y), r = std::hypot(ax, ay);
    if (r == 0) { c = 1; sn = 0; return; }
    if (ax == 0) { c = 0; sn = std::conj(y) / ay; return; }
    c = ax / r;
    sn = (x / ax) * std::conj(y) / r;
  };
  int hi = m - 1, its = 0, total = 0;
  while (hi > 0) {
    int lo = hi;
    for (; lo > 0; --lo) {
      double sc = std::abs(A(lo - 1, lo - 1)) + std::abs(A(lo, lo));
      if (sc == 0) sc = amax;
      if (std::abs(A(lo, lo - 1)) <= eps * sc) {
        A(lo, lo - 1) = 0;
        break;
      }
    }
    if (lo == hi) {
      --hi;
      its = 0;
      continue;
    }
    PP_REQUIRE(++total <= 60 * m, "eigenvector exponential: the QR iteration did not converge");
    ++its;
    cd mu;
    if (its % 11 == 0) {
      mu = A(hi, hi) + 0.75 * std::abs(A(hi, hi - 1));
    } else {
      const cd half = 0.5 * (A(hi - 1, hi - 1) - A(hi, hi));
      const cd disc = std::sqrt(half * half + A(hi - 1, hi) * A(hi, hi - 1));
      const cd m1 = A(hi, hi) + half + disc, m2 = A(hi, hi) + half - disc;
      mu = std::abs(m1 - A(hi, hi)) < std::abs(m2 - A(hi, hi)) ? m1 : m2;
    }
    for (int k = lo; k < hi; ++k) {
      double c;
      cd sn;
      if (k == lo) givens(A(k, k) - mu, A(k + 1, k), c, sn);
      else givens(A(k, k - 1), A(k + 1, k - 1), c, sn);
      rot(k, c, sn, k == lo ? lo : k - 1, std::min(hi, k + 2));
      if (k > lo) A(k + 1, k - 1) = 0;
    }
  }
  double tmax = 0;
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j) tmax = std::max(tmax, std::abs(A(i, j)));
  const double smin = std::max(eps * tmax, std::numeric_limits<double>::min());
  std::vector<cd> out(m, 0.0), y(m), v(m);
  for (int k = 0; k < m; ++k) {
    const cd wk = A(k, k);
    std::fill(y.begin(), y.end(), cd(0));
    y[k] = 1;
    for (int j = k - 1; j >= 0; --j) {
      cd acc = 0;
      for (int i = j + 1; i <= k; ++i) acc += A(j, i) * y[i];
      cd d = A(j, j) - wk;
      if (std::abs(d) < smin) d = smin;
      y[j] = -acc / d;
    }
    double nrm = 0;
    for (int i = 0; i < m; ++i) {
      cd t = 0;
      for (int j = 0; j <= k; ++j) t += q[(size_t)i * m + j] * y[j];
      v[i] = t;
      nrm += std::norm(t);
    }
    nrm = std::sqrt(nrm);
    // column 0 of V diag(exp(w)) V^H: sum_k v_k exp(w_k) conj(v_k[0])
    const cd coef = std::exp(wk) * std::conj(v[0]) / (nrm * nrm);
    for (int i = 0; i < m; ++i) out[i] += v[i] * coef;
  }
  return out;
}

Pamp::Pamp(long long n_, int m_, int order_, bool dorth_) : n(n_), m(m_), order(order_), dorth(dorth_) {
  PP_REQUIRE(m >= 2 && m <= 127, "pAMP basis size must be 2..127");
  PP_REQUIRE(order >= 1 && order < 30, "Pade order must be 1..29");
  V.resize((size_t)m * n);
  r.resize(n);
  H.resize((size_t)m * m);
  sc.resize(8);
  hbuf.resize(2 * (m + 4));
  expcol.resize(m);
  flags.resize(4);
  work.resize((size_t)6 * m * m);
  piv.resize(m);
}

void Pamp::propagate(pp_operator_fn op, void* ctx, double2* v, double2 dt, hipStream_t s) {
  auto apply_op = [&](const double2* in, double2* out) {
    ++op_count;
    int rc = op(ctx, reinterpret_cast<const pp_cplx*>(in), reinterpret_cast<pp_cplx*>(out), (pp_stream)s);
    if (rc != 0) throw Error(PP_ERR_INTERNAL, "pAMP operator callback failed");
  };
  op_count = 0;
  // Residual = input / |input|
  vec_norm2(n, v, sc.ptr, red, s);
  hipLaunchKernelGGL(pamp_init_kernel, dim3(1), dim3(256), 0, s, sc.ptr, H.ptr, m, flags.ptr);
  PP_HIP(hipGetLastError());
  vec_zero((long long)m * n, V.ptr, s);
  vec_scale_copy(n, v, r.ptr, sc.ptr + 1, false, s);
  // PerformInitialArnoldiStep (pamp.h:270-290)
  vec_copy(n, r.ptr, V.ptr, s);
  apply_op(V.ptr, r.ptr);
  vec_dot(n, V.ptr, r.ptr, H.ptr, red, s);          // H(0,0) = <v0, r>
  vec_axpy_dev(n, H.ptr, true, V.ptr, r.ptr, s);
  const int grid = (int)std::min<long long>(4096, (n + 255) / 256);
  // RecreateArnoldiFactorization: m-1 Arnoldi steps (pamp.h:419-427, 292-351)
  for (int j = 0; j < m - 1; ++j) {
    vec_norm2(n, r.ptr, hbuf.ptr + (m + 2), red, s);
    hipLaunchKernelGGL(pamp_beta_kernel, dim3(1), dim3(1), 0, s, hbuf.ptr + (m + 2), sc.ptr, H.ptr, m, j,
                       tol, flags.ptr);
    double2* vj1 = V.ptr + (size_t)(j + 1) * n;
    hipLaunchKernelGGL(pamp_newvec_kernel, dim3(grid), dim3(256), 0, s, n, r.ptr, vj1, sc.ptr, flags.ptr);
    PP_HIP(hipGetLastError());
    apply_op(vj1, r.ptr);
    const int k = j + 2;
    block_dot(n, k, V.ptr, n, r.ptr, hbuf.ptr, red, s);                       // h, |r|^2 (orig)
    block_update(n, k, V.ptr, n, hbuf.ptr, r.ptr, hbuf.ptr + (k + 1), red, s);  // r -= V h, |r|^2
    hipLaunchKernelGGL(pamp_store_h_kernel, dim3(1), dim3(128), 0, s, hbuf.ptr, H.ptr, m, j, k, flags.ptr, 0);
    if (dorth) {
      // PerformOrthogonalization (pamp.h:353-401): at most 3 conditional passes
      // hbuf[k] holds |r|^2 before the last projection ("orig"), hbuf[k+1] after it;
      // a pass recomputes both, so the next check compares the right pair.
      for (int pass = 0; pass < max_orth; ++pass) {
        hipLaunchKernelGGL(pamp_orth_check_kernel, dim3(1), dim3(1), 0, s, hbuf.ptr + k, hbuf.ptr + (k + 1), flags.ptr);
        block_dot_if(n, k, V.ptr, n, r.ptr, hbuf.ptr, red, flags.ptr + 1, s);
        block_update_if(n, k, V.ptr, n, hbuf.ptr, r.ptr, hbuf.ptr + (k + 1), red, flags.ptr + 1, s);
        hipLaunchKernelGGL(pamp_store_h_kernel, dim3(1), dim3(128), 0, s, hbuf.ptr, H.ptr, m, j, k, flags.ptr + 1, 2);
      }
    }
    PP_HIP(hipGetLastError());
  }
  if (method == 1) {  // ExponentiationEigenvector (pamp.h:640-643): on the host, one synchronisation
    std::vector<double2> hh((size_t)m * m);
    PP_HIP(hipMemcpyAsync(hh.data(), H.ptr, hh.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    const std::vector<std::complex<double>> col = eigenvector_exp_col0(m, hh, dt);
    std::vector<double2> c(m);
    for (int i = 0; i < m; ++i) c[i] = make_double2(col[i].real(), col[i].imag());
    PP_HIP(hipMemcpyAsync(expcol.ptr, c.data(), m * sizeof(double2), hipMemcpyHostToDevice, s));
    PP_HIP(hipMemcpyAsync(sc.ptr + 2, &c[m - 1], sizeof(double2), hipMemcpyHostToDevice, s));
    PP_HIP(hipStreamSynchronize(s));  // the host vector is released on return
  } else {
    hipLaunchKernelGGL(pade_exp_kernel, dim3(1), dim3(256), 0, s, H.ptr, m, dt, order, work.ptr, expcol.ptr, sc.ptr,
                       piv.ptr);
  }
  PP_HIP(hipGetLastError());
  // input = norm * V exp(H)[:, 0]
  vec_lincomb(n, m, V.ptr, n, expcol.ptr, sc.ptr, v, s);
}

// ----------------------------------------------------------------------------------
// Cayley
// ----------------------------------------------------------------------------------
int Cayley::operator_cb(void* ctx, const pp_cplx* in_, pp_cplx* out_, pp_stream s_) {
  auto* c = static_cast<Cayley*>(ctx);
  try {
    const double2* in = reinterpret_cast<const double2*>(in_);
    double2* out = reinterpret_cast<double2*>(out_);
    hipStream_t s = (hipStream_t)s_;
    // SolverCallback: MultiplySpH(+i dt/2) then left preconditioner (CayleyPropagator.py:145-150)
    vec_copy(c->n, in, c->tmp_s.ptr, s);
    c->ham->repr->overlap(PP_OVERLAP_MULTIPLY, false, c->tmp_s.ptr, s);
    if (c->fham) fham_apply64(*c->fham, c->scalings.data(), in, out, c->tmp_s.ptr, make_double2(0.0, 0.5 * c->dt), s);
    else c->ham->apply(c->scalings.data(), in, out, c->tmp_s.ptr, make_double2(0.0, 0.5 * c->dt), s);
    if (c->pc) precond_solve(*c->pc, out, s);
    else if (c->kron) kron_solve(*c->kron, make_double2(0.0, 0.5 * c->dt), out, s);
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
}

void Cayley::step(double2* psi, double dt_, const double* sc, hipStream_t s) {
  dt = dt_;
  scalings.assign(sc, sc + ham->terms.size());
  // TempPsi = (S - i dt/2 H) psi   (AdvanceStep, CayleyPropagator.py:81-85)
  vec_copy(n, psi, tmp_s.ptr, s);
  ham->repr->overlap(PP_OVERLAP_MULTIPLY, false, tmp_s.ptr, s);
  if (fham) fham_apply64(*fham, scalings.data(), psi, rhs.ptr, tmp_s.ptr, make_double2(0.0, -0.5 * dt), s);
  else ham->apply(scalings.data(), psi, rhs.ptr, tmp_s.ptr, make_double2(0.0, -0.5 * dt), s);
  if (pc) precond_solve(*pc, rhs.ptr, s);
  else if (kron) kron_solve(*kron, make_double2(0.0, 0.5 * dt), rhs.ptr, s);
  gmres.solve(&Cayley::operator_cb, this, rhs.ptr, psi, s);
}

// pAMP balanced-overlap step context
struct BalancedCtx {
  Hamiltonian* ham;
  const double* scalings;
};

static int balanced_cb(void* ctx, const pp_cplx* in, pp_cplx* out, pp_stream s) {
  auto* b = static_cast<BalancedCtx*>(ctx);
  try {
    b->ham->multiply_balanced(b->scalings, reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out),
                              (hipStream_t)s);
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
}

void pamp_step_balanced(Pamp& p, Hamiltonian& ham, double2* psi, double2 dt, const double* scalings, hipStream_t s) {
  // PamPropagator.AdvanceStep (PamPropagator.py:50-55)
  ham.repr->overlap(PP_OVERLAP_SQRT_MULTIPLY, false, psi, s);
  BalancedCtx ctx{&ham, scalings};
  p.propagate(&balanced_cb, &ctx, psi, dt, s);
  ham.repr->overlap(PP_OVERLAP_SQRT_SOLVE, false, psi, s);
}

}  // namespace pp

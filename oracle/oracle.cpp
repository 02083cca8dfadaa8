// ORACLE -- test infrastructure only.
//
// A sequential CPU restatement of PyProp's propagation hot path, used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker and the
// CPU baseline.  It is never linked into, loaded by or called from the product
// library (pyprop_amd/).
//
// Each function follows the cited reference code (paths relative to the PyProp tree):
//   oracle_tpmul             core/tensorpotential/tensorpotentialmultiply_generator.py:302-762,1227-1303
//                            (generated loops; innermost zhbmv/zgbmv as reference BLAS loops)
//   oracle_overlap_factor    core/representation/overlapmatrix.cpp:9-37 (zpbtrf 'U', unblocked)
//   oracle_overlap_apply     core/representation/overlapmatrix.cpp:111-229 (ztbmv/ztbsv loops)
//   oracle_gmres_solve       core/krylov/gmres/gmres.h:260-505 (+ zgels by Householder QR, zlarfg convention)
//   oracle_pamp_propagate    core/krylov/pamp/pamp/pamp.h:270-670 (Pade via zgetrf/zgetrs loops;
//                            EigenvectorExponential :567-608 via zgeev restated for Hessenberg input)
//   oracle_bspline_*         core/transform/bspline/bspline.cpp:26-161 (value/derivative
//                            recursions), :372-466 (BSplineGlobalOverlapIntegral),
//                            core/tensorpotential/basis_bspline.h:10-47 (one rank of
//                            RepresentPotentialInBasisBSpline), pyprop/bspline/BSpline.py:190-242
//                            (global quadrature grid)
//   oracle_sys_*             pyprop/tensorpotential/BasisPropagator.py:116-159,
//                            pyprop/propagator/CayleyPropagator.py:71-150,
//                            pyprop/propagator/PamPropagator.py:50-75,
//                            examples/tensor/helium_stabilization/preconditioner.py:113-198 (exact solve)
//
// oracle_set_threads(n) splits the multiply by rank-0 destination row, the overlap passes by
// fibre and the preconditioner solve by channel; each output element keeps the sequential
// summation order, so the results do not depend on n.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using cd = std::complex<double>;

enum { SIMP = 0, HERM = 1, DIAG = 2, IDENT = 3, BAND = 4, BANDNH = 5, DENSE = 6 };

static thread_local std::string g_err;

extern "C" const char* oracle_last_error() { return g_err.c_str(); }

template <class F>
static int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Host threads for the multiply, the overlap passes and the preconditioner solve.  Work is
// split so that every output element is computed by one thread in the sequential order,
// so any thread count gives bitwise the same result as one thread.
static int g_threads = 1;

extern "C" void oracle_set_threads(int n) { g_threads = std::max(1, std::min(n, 256)); }
extern "C" int oracle_get_threads() { return g_threads; }

// f(part, nparts) on g_threads threads (inline when 1); the first exception is rethrown
template <class F>
static void run_parts(F&& f) {
  const int nt = g_threads;
  if (nt <= 1) {
    f(0, 1);
    return;
  }
  std::vector<std::thread> pool;
  std::vector<std::string> errs(nt);
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      try {
        f(t, nt);
      } catch (const std::exception& e) {
        errs[t] = e.what();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (!e.empty()) throw std::runtime_error(e);
}

// ------------------------------------------------------------------------------------
// tensor potential multiply
// ------------------------------------------------------------------------------------
namespace {

struct TP {
  int R;
  const int* ids;
  const cd* V;
  const int* pext;
  double scaling;
  const cd* src;
  const int* sext;
  cd* dst;
  const int32_t* const* pairs;
  int part = 0, nparts = 1;  // this thread computes the rank-0 destination rows == part (mod nparts)
};

// reference zgbmv('N', m=n, n, kl=b, ku=b, alpha, A, lda=2b+1, x, 1, beta=1, y, 1)
void zgbmv_n(int n, int b, cd alpha, const cd* A, const cd* x, cd* y) {
  const int lda = 2 * b + 1;
  for (int j = 0; j < n; ++j) {
    if (x[j] == cd(0)) continue;
    const cd temp = alpha * x[j];
    const int k = b - j;
    for (int i = std::max(0, j - b); i <= std::min(n - 1, j + b); ++i) y[i] += temp * A[k + i + (size_t)j * lda];
  }
}

// reference zhbmv('L', n, k, alpha, A, lda=k+1, x, 1, beta=1, y, 1)
void zhbmv_l(int n, int k, cd alpha, const cd* A, const cd* x, cd* y) {
  const int lda = k + 1;
  for (int j = 0; j < n; ++j) {
    const cd temp1 = alpha * x[j];
    cd temp2 = 0;
    y[j] += temp1 * A[(size_t)j * lda].real();
    for (int i = j + 1; i <= std::min(n - 1, j + k); ++i) {
      y[i] += temp1 * A[(i - j) + (size_t)j * lda];
      temp2 += std::conj(A[(i - j) + (size_t)j * lda]) * x[i];
    }
    y[j] += alpha * temp2;
  }
}

void innermost_entry(const TP& t, bool conj, size_t dOff, size_t sOff, size_t pOff, int row, int col, int i) {
  const cd v = conj ? std::conj(t.V[pOff + i]) : t.V[pOff + i];
  t.dst[dOff + row] = t.dst[dOff + row] + (v * t.scaling) * t.src[sOff + col];
}

void loop(const TP& t, int r, bool conj, size_t dOff, size_t sOff, size_t pOff) {
  const int id = t.ids[r];
  const int P = t.pext[r];
  const int N = t.sext[r];
  const bool inner = (r == t.R - 1);
  auto rec = [&](int row, int col, int i, bool cj) {
    if (r == 0 && t.nparts > 1 && row % t.nparts != t.part) return;
    if (inner) innermost_entry(t, cj, dOff * N, sOff * N, pOff * P, row, col, i);
    else loop(t, r + 1, cj, dOff * N + row, sOff * N + col, pOff * P + i);
  };
  switch (id) {
    case SIMP:
      for (int i = 0; i < P; ++i) rec(t.pairs[r][2 * i], t.pairs[r][2 * i + 1], i, conj);
      break;
    case HERM:
      for (int i = 0; i < P; ++i) {
        rec(t.pairs[r][2 * i], t.pairs[r][2 * i + 1], i, conj);
        rec(t.pairs[r][2 * i + 1], t.pairs[r][2 * i], i, !conj);
      }
      break;
    case DIAG:
      for (int i = 0; i < P; ++i) rec(i, i, i, conj);
      break;
    case IDENT:
      for (int row = 0; row < N; ++row) rec(row, row, 0, conj);
      break;
    case DENSE:
      for (int row = 0; row < N; ++row)
        for (int col = 0; col < N; ++col) rec(row, col, row * N + col, conj);
      break;
    case BAND: {
      const int band = P / N;
      if (inner) {
        zhbmv_l(N, band - 1, cd(t.scaling, 0), t.V + pOff * P, t.src + sOff * N, t.dst + dOff * N);
        break;
      }
      for (int row = 0; row < N; ++row)
        for (int col = row; col < std::min(row + band, N); ++col) rec(row, col, row * band + col - row, conj);
      for (int hr = 0; hr < N; ++hr)
        for (int hc = hr + 1; hc < std::min(hr + band, N); ++hc) rec(hc, hr, hr * band + hc - hr, !conj);
      break;
    }
    case BANDNH: {
      const int bw = P / N;
      const int b = (bw - 1) / 2;
      if (inner && !conj) {
        zgbmv_n(N, b, cd(t.scaling, 0), t.V + pOff * P, t.src + sOff * N, t.dst + dOff * N);
        break;
      }
      for (int col = 0; col < N; ++col)
        for (int row = std::max(0, col - b); row <= std::min(N - 1, col + b); ++row)
          rec(row, col, col * bw + b + row - col, conj);
      break;
    }
    default:
      throw std::runtime_error("unknown storage id");
  }
}

}  // namespace

extern "C" int oracle_tpmul(int rank, const int* ids, const double* potential, const int* pot_ext, double scaling,
                            const double* src, const int* src_ext, double* dst, const int32_t* const* pairs) {
  return guard([&] {
    TP t{rank, ids, reinterpret_cast<const cd*>(potential), pot_ext, scaling, reinterpret_cast<const cd*>(src),
         src_ext, reinterpret_cast<cd*>(dst), pairs};
    if (rank < 2) {  // one fibre: the innermost BLAS call is not split
      loop(t, 0, false, 0, 0, 0);
      return;
    }
    run_parts([&](int part, int nparts) {
      TP tt = t;
      tt.part = part;
      tt.nparts = nparts;
      loop(tt, 0, false, 0, 0, 0);
    });
  });
}

// ------------------------------------------------------------------------------------
// overlap
// ------------------------------------------------------------------------------------
// zpbtf2('U') on a real band: for each j, U(j,j) = sqrt(a_jj), row j of the band scaled
// by 1/U(j,j), trailing rank-1 update (zher).  Storage [j*(sd+1) + sd + i - j].
extern "C" int oracle_overlap_factor(int n, int sd, const double* s_upper, double* u) {
  return guard([&] {
    const int ld = sd + 1;
    std::copy(s_upper, s_upper + (size_t)n * ld, u);
    auto A = [&](int i, int j) -> double& { return u[(size_t)j * ld + sd + i - j]; };
    for (int j = 0; j < n; ++j) {
      double ajj = A(j, j);
      if (ajj <= 0) throw std::runtime_error("overlap not positive definite");
      ajj = std::sqrt(ajj);
      A(j, j) = ajj;
      const int kn = std::min(sd, n - 1 - j);
      if (kn > 0) {
        for (int c = 1; c <= kn; ++c) A(j, j + c) /= ajj;
        for (int c1 = 1; c1 <= kn; ++c1)
          for (int c2 = c1; c2 <= kn; ++c2) A(j + c1, j + c2) -= A(j, j + c1) * A(j, j + c2);
      }
    }
  });
}

namespace {
void tb_fibre(int n, int sd, const double* u, int op, bool conj, cd* x, size_t st) {
  auto U = [&](int i, int j) { return u[(size_t)j * (sd + 1) + sd + i - j]; };
  auto X = [&](int i) -> cd& { return x[i * st]; };
  // ztbmv('U','N'): for j ascending: temp = x_j; x_i += temp*U(i,j) for i < j; x_j *= U(j,j)
  auto mul_u = [&] {
    for (int j = 0; j < n; ++j) {
      if (X(j) == cd(0)) continue;
      const cd temp = X(j);
      for (int i = std::max(0, j - sd); i < j; ++i) X(i) += temp * U(i, j);
      X(j) *= U(j, j);
    }
  };
  // ztbmv('U','C'): for j descending: temp = x_j * U(j,j) + sum_{i<j} U(i,j) x_i
  auto mul_uh = [&] {
    for (int j = n - 1; j >= 0; --j) {
      cd temp = X(j) * U(j, j);
      for (int i = j - 1; i >= std::max(0, j - sd); --i) temp += U(i, j) * X(i);
      X(j) = temp;
    }
  };
  // ztbsv('U','N'): for j descending: x_j /= U(j,j); x_i -= x_j U(i,j)
  auto solve_u = [&] {
    for (int j = n - 1; j >= 0; --j) {
      if (X(j) == cd(0)) continue;
      X(j) /= U(j, j);
      const cd temp = X(j);
      for (int i = j - 1; i >= std::max(0, j - sd); --i) X(i) -= temp * U(i, j);
    }
  };
  // ztbsv('U','C'): for j ascending: temp = x_j - sum U(i,j) x_i; x_j = temp / U(j,j)
  auto solve_uh = [&] {
    for (int j = 0; j < n; ++j) {
      cd temp = X(j);
      for (int i = std::max(0, j - sd); i < j; ++i) temp -= U(i, j) * X(i);
      X(j) = temp / U(j, j);
    }
  };
  if (op == 0) { mul_u(); mul_uh(); }
  else if (op == 1) { solve_uh(); solve_u(); }
  else if (op == 2) { if (conj) mul_uh(); else mul_u(); }
  else { if (conj) solve_uh(); else solve_u(); }
}
}  // namespace

extern "C" int oracle_overlap_apply(int n, int sd, const double* u, int op, int conj, int ndim, const int* ext,
                                    int rank, double* data) {
  return guard([&] {
    size_t pre = 1, post = 1;
    for (int r = 0; r < rank; ++r) pre *= ext[r];
    for (int r = rank + 1; r < ndim; ++r) post *= ext[r];
    if (ext[rank] != n) throw std::runtime_error("overlap size mismatch");
    cd* x = reinterpret_cast<cd*>(data);
    run_parts([&](int part, int nparts) {
      for (size_t f = part; f < pre * post; f += nparts) {
        const size_t p = f / post, q = f % post;
        tb_fibre(n, sd, u, op, conj != 0, x + p * n * post + q, post);
      }
    });
  });
}

// ------------------------------------------------------------------------------------
// BLAS-1 as the reference's BLASIntegrationFunctor uses them
// ------------------------------------------------------------------------------------
namespace {
cd zdotc(size_t n, const cd* x, const cd* y) {
  cd s = 0;
  for (size_t i = 0; i < n; ++i) s += std::conj(x[i]) * y[i];
  return s;
}
double vnorm(size_t n, const cd* x) { return std::sqrt(zdotc(n, x, x).real()); }
void zaxpy(size_t n, cd a, const cd* x, cd* y) {
  for (size_t i = 0; i < n; ++i) y[i] += a * x[i];
}
void zscal(size_t n, cd a, cd* x) {
  for (size_t i = 0; i < n; ++i) x[i] *= a;
}
// temp = V^H r over k vectors (zgemv 'C')
void block_dot(size_t n, int k, const cd* V, const cd* r, cd* out) {
  for (int i = 0; i < k; ++i) out[i] = zdotc(n, V + (size_t)i * n, r);
}
// t = V h (zgemv 'N')
void block_mv(size_t n, int k, const cd* V, const cd* h, cd* t) {
  std::fill(t, t + n, cd(0));
  for (int i = 0; i < k; ++i) zaxpy(n, h[i], V + (size_t)i * n, t);
}

// LAPACK zlarfg: H (beta; 0) = (alpha; x), H = I - tau v v^H, v(0) = 1
void zlarfg(int n, cd& alpha, cd* x, cd& tau) {
  if (n <= 0) { tau = 0; return; }
  double xnorm = 0;
  for (int i = 0; i < n - 1; ++i) xnorm += std::norm(x[i]);
  xnorm = std::sqrt(xnorm);
  const double ar = alpha.real(), ai = alpha.imag();
  if (xnorm == 0 && ai == 0) { tau = 0; return; }
  double beta = -std::copysign(std::sqrt(ar * ar + ai * ai + xnorm * xnorm), ar);
  tau = cd((beta - ar) / beta, -ai / beta);
  const cd scal = 1.0 / (alpha - beta);
  for (int i = 0; i < n - 1; ++i) x[i] *= scal;
  alpha = beta;
}

// zgels('N') for rows >= cols: A column-major (rows x cols), b (rows) -> x (cols)
void zgels(int rows, int cols, std::vector<cd> A, std::vector<cd> b, std::vector<cd>& x) {
  for (int k = 0; k < cols; ++k) {
    cd tau;
    cd alpha = A[(size_t)k * rows + k];
    zlarfg(rows - k, alpha, &A[(size_t)k * rows + k + 1], tau);
    A[(size_t)k * rows + k] = alpha;
    // apply H^H = I - conj(tau) v v^H to trailing columns and to b (zunmqr 'C')
    auto apply = [&](cd* col) {
      cd w = col[k];
      for (int i = k + 1; i < rows; ++i) w += std::conj(A[(size_t)k * rows + i]) * col[i];
      w *= std::conj(tau);
      col[k] -= w;
      for (int i = k + 1; i < rows; ++i) col[i] -= w * A[(size_t)k * rows + i];
    };
    for (int c = k + 1; c < cols; ++c) apply(&A[(size_t)c * rows]);
    apply(b.data());
  }
  x.assign(cols, 0);
  for (int k = cols - 1; k >= 0; --k) {
    cd s = b[k];
    for (int c = k + 1; c < cols; ++c) s -= A[(size_t)c * rows + k] * x[c];
    x[k] = s / A[(size_t)k * rows + k];
  }
}
}  // namespace

typedef int (*oracle_op)(void* ctx, const double* in, double* out);

// ------------------------------------------------------------------------------------
// GMRES
// ------------------------------------------------------------------------------------
struct GmresResult {
  double error;
  int op_count;
  int orth_count;
  int steps;
};

static GmresResult gmres_solve(size_t N, int m, double tol, bool dorth, oracle_op op, void* ctx, const cd* rhs,
                               cd* sol, double* err_list) {
  if (tol < 1e-13) tol = 1e-13;
  const int maxOrth = 5;
  std::vector<cd> V((size_t)m * N, 0), r(rhs, rhs + N), tmp(N);
  std::vector<cd> H((size_t)m * (m + 1), 0);  // H(col,row) at col*(m+1)+row
  auto Hc = [&](int col, int row) -> cd& { return H[(size_t)col * (m + 1) + row]; };
  std::vector<cd> sol_coeff(m + 1, 0);
  GmresResult res{0, 0, 0, 0};
  if (err_list) std::fill(err_list, err_list + m, 0.0);
  auto apply = [&](const cd* in, cd* out) {
    ++res.op_count;
    if (op(ctx, reinterpret_cast<const double*>(in), reinterpret_cast<double*>(out)) != 0)
      throw std::runtime_error("operator failed");
  };
  const double inputNorm = vnorm(N, r.data());
  zscal(N, 1.0 / inputNorm, r.data());
  // initial Arnoldi step
  std::copy(r.begin(), r.end(), V.begin());
  apply(V.data(), r.data());
  cd alpha = zdotc(N, V.data(), r.data());
  zaxpy(N, -alpha, V.data(), r.data());
  Hc(0, 0) = alpha;
  double beta = vnorm(N, r.data());
  bool happy = std::abs(beta) < tol;
  Hc(0, 1) = beta;
  if (happy) sol_coeff[0] = inputNorm / Hc(0, 0);
  int step = 0;
  double err = std::abs(Hc(step, step + 1));
  if (err_list) err_list[step] = err;
  std::vector<cd> ov(m + 1), ov3(m + 1);
  while (step < m - 1 && !happy && err > tol) {
    const int j = step;
    cd* vj1 = V.data() + (size_t)(j + 1) * N;
    std::copy(r.begin(), r.end(), vj1);
    zscal(N, 1.0 / Hc(j, j + 1), vj1);
    apply(vj1, r.data());
    double origNorm = vnorm(N, r.data());
    block_dot(N, j + 2, V.data(), r.data(), ov.data());
    block_mv(N, j + 2, V.data(), ov.data(), tmp.data());
    zaxpy(N, -1.0, tmp.data(), r.data());
    double residualNorm = vnorm(N, r.data());
    if (dorth) {
      int cur = 0;
      while (std::abs(residualNorm) < 0.717 * std::abs(origNorm)) {
        ++res.orth_count;
        ++cur;
        if (cur > maxOrth) break;
        block_dot(N, j + 2, V.data(), r.data(), ov3.data());
        block_mv(N, j + 2, V.data(), ov3.data(), tmp.data());
        zaxpy(N, -1.0, tmp.data(), r.data());
        for (int i = 0; i < j + 2; ++i) ov[i] += ov3[i];
        origNorm = residualNorm;
        residualNorm = vnorm(N, r.data());
      }
    }
    for (int i = 0; i < j + 2; ++i) Hc(j + 1, i) = ov[i];
    beta = vnorm(N, r.data());
    if (std::abs(beta) < tol) happy = true;
    Hc(j + 1, j + 2) = beta;
    ++step;
    const int ks = step + 1;
    std::vector<cd> A((size_t)(ks + 1) * ks), b(ks + 1, 0), x;
    for (int c = 0; c < ks; ++c)
      for (int rr = 0; rr <= ks; ++rr) A[(size_t)c * (ks + 1) + rr] = Hc(c, rr);
    b[0] = inputNorm;
    zgels(ks + 1, ks, A, b, x);
    for (int i = 0; i < ks; ++i) sol_coeff[i] = x[i];
    std::vector<cd> rv(ks + 1, 0);
    for (int rr = 0; rr <= ks; ++rr)
      for (int c = 0; c < ks; ++c) rv[rr] += Hc(c, rr) * x[c];
    rv[0] -= inputNorm;
    err = vnorm(ks + 1, rv.data());
    if (err_list) err_list[step] = err;
  }
  std::fill(sol, sol + N, cd(0));
  for (int i = 0; i <= step; ++i) zaxpy(N, sol_coeff[i], V.data() + (size_t)i * N, sol);
  res.error = err;
  res.steps = step;
  return res;
}

extern "C" int oracle_gmres_solve(int64_t n, int m, double tol, int dorth, oracle_op op, void* ctx,
                                  const double* rhs, double* sol, double* out_error, int* out_ops, double* err_list) {
  return guard([&] {
    GmresResult r = gmres_solve((size_t)n, m, tol, dorth != 0, op, ctx, reinterpret_cast<const cd*>(rhs),
                                reinterpret_cast<cd*>(sol), err_list);
    if (out_error) *out_error = r.error;
    if (out_ops) *out_ops = r.op_count;
  });
}

// ------------------------------------------------------------------------------------
// pAMP
// ------------------------------------------------------------------------------------
namespace {
// C = A B, column-major m x m
void mm(int m, const cd* A, const cd* B, cd* C) {
  for (int col = 0; col < m; ++col)
    for (int row = 0; row < m; ++row) {
      cd s = 0;
      for (int k = 0; k < m; ++k) s += A[(size_t)k * m + row] * B[(size_t)col * m + k];
      C[(size_t)col * m + row] = s;
    }
}

// zgetrf (partial pivoting by |re|+|im|) and zgetrs('N') with m right-hand sides, column-major
void lu_solve(int m, std::vector<cd> A, cd* B) {
  std::vector<int> piv(m);
  for (int k = 0; k < m; ++k) {
    int p = k;
    double best = -1;
    for (int i = k; i < m; ++i) {
      const cd v = A[(size_t)k * m + i];
      const double a = std::abs(v.real()) + std::abs(v.imag());
      if (a > best) { best = a; p = i; }
    }
    piv[k] = p;
    if (A[(size_t)k * m + p] == cd(0)) throw std::runtime_error("singular Pade denominator");
    if (p != k)
      for (int c = 0; c < m; ++c) std::swap(A[(size_t)c * m + k], A[(size_t)c * m + p]);
    const cd inv = 1.0 / A[(size_t)k * m + k];
    for (int i = k + 1; i < m; ++i) A[(size_t)k * m + i] *= inv;
    for (int c = k + 1; c < m; ++c) {
      const cd t = A[(size_t)c * m + k];
      for (int i = k + 1; i < m; ++i) A[(size_t)c * m + i] -= A[(size_t)k * m + i] * t;
    }
  }
  for (int col = 0; col < m; ++col) {
    cd* b = B + (size_t)col * m;
    for (int k = 0; k < m; ++k)
      if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int i = 0; i < m; ++i)
      for (int k = 0; k < i; ++k) b[i] -= A[(size_t)k * m + i] * b[k];
    for (int i = m - 1; i >= 0; --i) {
      for (int k = i + 1; k < m; ++k) b[i] -= A[(size_t)k * m + i] * b[k];
      b[i] /= A[(size_t)i * m + i];
    }
  }
}

void pade_exp(int m, std::vector<cd> A, int order, std::vector<cd>& out) {
  // ScalingPadeExponential (pamp.h:430-468)
  double mx = 0;
  for (auto& z : A) mx = std::max(mx, std::abs(z));
  int s = (mx > 0) ? (int)std::ceil(std::log2(mx)) : 0;
  s = std::max(0, s);
  const double div = std::pow(2.0, s);
  for (auto& z : A) z /= div;
  // PadeExponential (pamp.h:470-565)
  const size_t mm2 = (size_t)m * m;
  std::vector<cd> A2(mm2), P(mm2, 0), Q(mm2, 0), T(mm2);
  mm(m, A.data(), A.data(), A2.data());
  std::vector<double> c(order + 1);
  c[0] = 1;
  for (int i = 0; i < order; ++i) c[i + 1] = c[i] * (double)(order - i) / (double)((i + 1) * (2 * order - i));
  for (int i = 0; i < m; ++i) {
    Q[(size_t)i * m + i] = c[order];
    P[(size_t)i * m + i] = c[order - 1];
  }
  bool odd = true;
  for (int i = order - 2; i >= 0; --i) {
    std::vector<cd>& X = odd ? Q : P;
    mm(m, X.data(), A2.data(), T.data());
    X.swap(T);
    for (int r = 0; r < m; ++r) X[(size_t)r * m + r] += c[i];
    odd = !odd;
  }
  {
    std::vector<cd>& X = odd ? Q : P;
    mm(m, X.data(), A.data(), T.data());
    X.swap(T);
  }
  for (size_t i = 0; i < mm2; ++i) Q[i] -= P[i];
  lu_solve(m, Q, P.data());
  out.assign(mm2, 0);
  for (size_t i = 0; i < mm2; ++i) out[i] = 2.0 * P[i];
  for (int i = 0; i < m; ++i) out[(size_t)i * m + i] += 1.0;
  if (odd)
    for (auto& z : out) z = -z;
  for (int i = 0; i < s; ++i) {
    mm(m, out.data(), out.data(), T.data());
    out.swap(T);
  }
}
// zgeev restated for the upper Hessenberg case: eigenvalues w and unit-norm eigenvectors
// (column k of V, row-major V[i*m + k]) of the m x m upper Hessenberg matrix a (row-major).
// Schur form by the single-shift QR algorithm (Wilkinson shift, an exceptional shift every 11th
// iteration, Givens bulge chase), eigenvectors of the triangular factor by back substitution
// (ztrevc, small divisors replaced by eps |T|), then v = Q y, normalised.
void hess_eig(int m, std::vector<cd> a, std::vector<cd>& w, std::vector<cd>& V) {
  auto A = [&](int i, int j) -> cd& { return a[(size_t)i * m + j]; };
  std::vector<cd> Q((size_t)m * m, 0);
  for (int i = 0; i < m; ++i) Q[(size_t)i * m + i] = 1;
  const double eps = std::numeric_limits<double>::epsilon();
  double anorm = 0;
  for (auto& z : a) anorm = std::max(anorm, std::abs(z));
  auto givens = [&](cd x, cd y, double& c, cd& sn) {
    const double ax = std::abs(x), ay = std::abs(y), r = std::hypot(ax, ay);
    if (r == 0) { c = 1; sn = 0; return; }
    if (ax == 0) { c = 0; sn = std::conj(y) / ay; return; }
    const cd al = x / ax;
    c = ax / r;
    sn = al * std::conj(y) / r;
  };
  // rows k, k+1 <- G rows, G = [c s; -conj(s) c]; columns k, k+1 <- columns G^H; Q <- Q G^H
  auto rotate = [&](int k, double c, cd sn, int col0, int row1) {
    for (int j = col0; j < m; ++j) {
      const cd x = A(k, j), y = A(k + 1, j);
      A(k, j) = c * x + sn * y;
      A(k + 1, j) = -std::conj(sn) * x + c * y;
    }
    for (int i = 0; i <= row1; ++i) {
      const cd x = A(i, k), y = A(i, k + 1);
      A(i, k) = c * x + std::conj(sn) * y;
      A(i, k + 1) = -sn * x + c * y;
    }
    for (int i = 0; i < m; ++i) {
      cd& x = Q[(size_t)i * m + k];
      cd& y = Q[(size_t)i * m + k + 1];
      const cd qx = x, qy = y;
      x = c * qx + std::conj(sn) * qy;
      y = -sn * qx + c * qy;
    }
  };
  int hi = m - 1, iter = 0, total = 0;
  while (hi > 0) {
    int lo = hi;
    while (lo > 0) {
      double sc = std::abs(A(lo - 1, lo - 1)) + std::abs(A(lo, lo));
      if (sc == 0) sc = anorm;
      if (std::abs(A(lo, lo - 1)) <= eps * sc) {
        A(lo, lo - 1) = 0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++total > 60 * m) throw std::runtime_error("hess_eig: QR iteration did not converge");
    ++iter;
    cd mu;
    if (iter % 11 == 0) {  // exceptional shift
      mu = A(hi, hi) + 0.75 * std::abs(A(hi, hi - 1));
    } else {  // the eigenvalue of the trailing 2 x 2 block closer to its last diagonal entry
      const cd p = A(hi - 1, hi - 1), q = A(hi - 1, hi), r = A(hi, hi - 1), t = A(hi, hi);
      const cd half = 0.5 * (p - t);
      const cd disc = std::sqrt(half * half + q * r);
      const cd m1 = t + half + disc, m2 = t + half - disc;
      mu = std::abs(m1 - t) < std::abs(m2 - t) ? m1 : m2;
    }
    for (int k = lo; k < hi; ++k) {
      double c;
      cd sn;
      if (k == lo) givens(A(k, k) - mu, A(k + 1, k), c, sn);
      else givens(A(k, k - 1), A(k + 1, k - 1), c, sn);
      rotate(k, c, sn, k == lo ? lo : k - 1, std::min(hi, k + 2));
      if (k > lo) A(k + 1, k - 1) = 0;
    }
  }
  w.resize(m);
  for (int i = 0; i < m; ++i) w[i] = A(i, i);
  double tnorm = 0;
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j) tnorm = std::max(tnorm, std::abs(A(i, j)));
  const double smin = std::max(eps * tnorm, std::numeric_limits<double>::min());
  V.assign((size_t)m * m, 0);
  std::vector<cd> y(m);
  for (int k = 0; k < m; ++k) {
    std::fill(y.begin(), y.end(), cd(0));
    y[k] = 1;
    for (int j = k - 1; j >= 0; --j) {
      cd acc = 0;
      for (int i = j + 1; i <= k; ++i) acc += A(j, i) * y[i];
      cd d = A(j, j) - w[k];
      if (std::abs(d) < smin) d = smin;
      y[j] = -acc / d;
    }
    double nrm = 0;
    for (int i = 0; i < m; ++i) {
      cd v = 0;
      for (int j = 0; j <= k; ++j) v += Q[(size_t)i * m + j] * y[j];
      V[(size_t)i * m + k] = v;
      nrm += std::norm(v);
    }
    nrm = std::sqrt(nrm);
    for (int i = 0; i < m; ++i) V[(size_t)i * m + k] /= nrm;
  }
}

// EigenvectorExponential (pamp.h:567-608): exp(A) ~ V diag(exp(w)) V^H with the unit-norm
// eigenvectors of A (exact for a normal A, as -i dt H_m is in exact arithmetic); A and out in
// the Hessenberg storage of pamp_propagate (element (row, col) at [col * m + row])
void eig_exp(int m, const std::vector<cd>& Acm, std::vector<cd>& out) {
  std::vector<cd> a((size_t)m * m);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) a[(size_t)r * m + c] = Acm[(size_t)c * m + r];
  std::vector<cd> w, V;
  hess_eig(m, a, w, V);
  out.assign((size_t)m * m, 0);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) {
      cd acc = 0;
      for (int k = 0; k < m; ++k) acc += V[(size_t)r * m + k] * std::exp(w[k]) * std::conj(V[(size_t)c * m + k]);
      out[(size_t)c * m + r] = acc;
    }
}
}  // namespace

extern "C" int oracle_pade_exp(int m, const double* A, int order, double* out) {
  return guard([&] {
    std::vector<cd> a(reinterpret_cast<const cd*>(A), reinterpret_cast<const cd*>(A) + (size_t)m * m), o;
    pade_exp(m, a, order, o);
    std::copy(o.begin(), o.end(), reinterpret_cast<cd*>(out));
  });
}

// V diag(exp(w)) V^H of an upper Hessenberg A, both in column-major storage
extern "C" int oracle_eig_exp(int m, const double* A, double* out) {
  return guard([&] {
    std::vector<cd> a(reinterpret_cast<const cd*>(A), reinterpret_cast<const cd*>(A) + (size_t)m * m), o;
    eig_exp(m, a, o);
    std::copy(o.begin(), o.end(), reinterpret_cast<cd*>(out));
  });
}

static cd pamp_propagate(size_t N, int m, int order, bool dorth, oracle_op op, void* ctx, cd* v, cd dt,
                         int* op_count, int* breakdown, int method = 0) {
  const double tol = std::sqrt(std::numeric_limits<double>::epsilon());
  const int maxOrth = 3;
  std::vector<cd> V((size_t)m * N, 0), r(v, v + N), tmp(N);
  std::vector<cd> H((size_t)m * m, 0);  // column-major: H(col,row) at col*m + row
  auto Hc = [&](int col, int row) -> cd& { return H[(size_t)col * m + row]; };
  int ops = 0;
  auto apply = [&](const cd* in, cd* out) {
    ++ops;
    if (op(ctx, reinterpret_cast<const double*>(in), reinterpret_cast<double*>(out)) != 0)
      throw std::runtime_error("operator failed");
  };
  const double norm = vnorm(N, r.data());
  zscal(N, 1.0 / norm, r.data());
  std::copy(r.begin(), r.end(), V.begin());
  apply(V.data(), r.data());
  cd alpha = zdotc(N, V.data(), r.data());
  zaxpy(N, -alpha, V.data(), r.data());
  Hc(0, 0) = alpha;
  bool happy = false;
  std::vector<cd> ov(m + 1), ov3(m + 1);
  for (int j = 0; j < m - 1 && !happy; ++j) {
    const double beta = vnorm(N, r.data());
    if (std::abs(beta) < tol) {
      happy = true;
      break;
    }
    zscal(N, 1.0 / beta, r.data());
    cd* vj1 = V.data() + (size_t)(j + 1) * N;
    std::copy(r.begin(), r.end(), vj1);
    Hc(j, j + 1) = beta;
    apply(vj1, r.data());
    double origNorm = vnorm(N, r.data());
    block_dot(N, j + 2, V.data(), r.data(), ov.data());
    block_mv(N, j + 2, V.data(), ov.data(), tmp.data());
    zaxpy(N, -1.0, tmp.data(), r.data());
    double residualNorm = vnorm(N, r.data());
    if (dorth) {
      int cur = 0;
      while (std::abs(residualNorm) < 0.717 * std::abs(origNorm)) {
        ++cur;
        if (cur > maxOrth) break;
        block_dot(N, j + 2, V.data(), r.data(), ov3.data());
        block_mv(N, j + 2, V.data(), ov3.data(), tmp.data());
        zaxpy(N, -1.0, tmp.data(), r.data());
        for (int i = 0; i < j + 2; ++i) ov[i] += ov3[i];
        origNorm = residualNorm;
        residualNorm = vnorm(N, r.data());
      }
    }
    for (int i = 0; i < j + 2; ++i) Hc(j + 1, i) = ov[i];
  }
  for (auto& z : H) z *= -dt * cd(0, 1);
  std::vector<cd> E;
  if (method == 1) eig_exp(m, H, E);  // ExponentiationEigenvector (pamp.h:640-643)
  else pade_exp(m, H, order, E);
  std::fill(v, v + N, cd(0));
  for (int i = 0; i < m; ++i) zaxpy(N, E[i], V.data() + (size_t)i * N, v);  // column 0
  zscal(N, norm, v);
  if (op_count) *op_count = ops;
  if (breakdown) *breakdown = happy ? 1 : 0;
  return E[m - 1];
}

extern "C" int oracle_pamp_propagate(int64_t n, int m, int order, int dorth, oracle_op op, void* ctx, double* v,
                                     double dt_re, double dt_im, double* err_est, int* ops, int* breakdown,
                                     int method) {
  return guard([&] {
    cd e = pamp_propagate((size_t)n, m, order, dorth != 0, op, ctx, reinterpret_cast<cd*>(v), cd(dt_re, dt_im), ops,
                          breakdown, method);
    if (err_est) {
      err_est[0] = e.real();
      err_est[1] = e.imag();
    }
  });
}

// ------------------------------------------------------------------------------------
// system: Hamiltonian (list of tensor potentials) + overlaps + preconditioner
// ------------------------------------------------------------------------------------
struct OracleTerm {
  int R;
  std::vector<int> ids, pext, sext;
  std::vector<cd> V;
  std::vector<std::vector<int32_t>> pairs;
};

struct OracleSystem {
  int ndim;
  std::vector<int> ext;
  size_t N;
  std::vector<OracleTerm> terms;
  struct Ov {
    int n = 0, sd = 0;
    std::vector<double> u;
  };
  std::vector<Ov> ov;
  // preconditioner: per channel banded LU without pivoting of the (n1*n2) block, band w
  int pc_nchan = 0, pc_n1 = 0, pc_n2 = 0, pc_w = 0;
  std::vector<cd> pc_lu;  // [c][i][w + j - i] for |i-j| <= w
  bool has_pc = false;
  // scratch
  std::vector<double> scal;
  double dt = 0;
  std::vector<cd> t1, t2;
};

extern "C" void* oracle_sys_create(int ndim, const int* ext) {
  auto* s = new OracleSystem();
  s->ndim = ndim;
  s->ext.assign(ext, ext + ndim);
  s->N = 1;
  for (int e : s->ext) s->N *= e;
  s->ov.resize(ndim);
  return s;
}

extern "C" void oracle_sys_destroy(void* p) { delete static_cast<OracleSystem*>(p); }

extern "C" int oracle_sys_add_potential(void* p, const int* ids, const double* V, const int* pext,
                                        const int32_t* const* pairs, const int* pair_counts) {
  return guard([&] {
    auto* s = static_cast<OracleSystem*>(p);
    OracleTerm t;
    t.R = s->ndim;
    t.ids.assign(ids, ids + t.R);
    t.pext.assign(pext, pext + t.R);
    t.sext = s->ext;
    size_t cnt = 1;
    for (int e : t.pext) cnt *= e;
    t.V.assign(reinterpret_cast<const cd*>(V), reinterpret_cast<const cd*>(V) + cnt);
    t.pairs.resize(t.R);
    for (int r = 0; r < t.R; ++r)
      if (pairs && pairs[r]) t.pairs[r].assign(pairs[r], pairs[r] + 2 * (size_t)pair_counts[r]);
    s->terms.push_back(std::move(t));
  });
}

extern "C" int oracle_sys_set_overlap(void* p, int rank, int n, int sd, const double* s_upper) {
  return guard([&] {
    auto* s = static_cast<OracleSystem*>(p);
    s->ov[rank].n = n;
    s->ov[rank].sd = sd;
    s->ov[rank].u.resize((size_t)n * (sd + 1));
    if (oracle_overlap_factor(n, sd, s_upper, s->ov[rank].u.data()) != 0) throw std::runtime_error(g_err);
  });
}

static void sys_overlap(OracleSystem* s, int op, bool conj, cd* psi) {
  for (int r = 0; r < s->ndim; ++r)
    if (s->ov[r].n)
      oracle_overlap_apply(s->ov[r].n, s->ov[r].sd, s->ov[r].u.data(), op, conj, s->ndim, s->ext.data(), r,
                           reinterpret_cast<double*>(psi));
}

// dst = sum_t scal[t] V_t src  (MultiplyHamiltonianNoOverlap)
static void sys_ham(OracleSystem* s, const double* scal, const cd* src, cd* dst) {
  std::fill(dst, dst + s->N, cd(0));
  for (size_t t = 0; t < s->terms.size(); ++t) {
    auto& T = s->terms[t];
    std::vector<const int32_t*> pp(T.R, nullptr);
    for (int r = 0; r < T.R; ++r) pp[r] = T.pairs[r].empty() ? nullptr : T.pairs[r].data();
    oracle_tpmul(T.R, T.ids.data(), reinterpret_cast<const double*>(T.V.data()), T.pext.data(), scal ? scal[t] : 1.0,
                 reinterpret_cast<const double*>(src), T.sext.data(), reinterpret_cast<double*>(dst), pp.data());
  }
}

extern "C" int oracle_sys_multiply(void* p, int mode, const double* scal, const double* src, double* dst) {
  // mode 0: no overlap; 1: S^-1 H; 2: balanced U^-H H U^-1
  return guard([&] {
    auto* s = static_cast<OracleSystem*>(p);
    const cd* in = reinterpret_cast<const cd*>(src);
    cd* out = reinterpret_cast<cd*>(dst);
    if (mode == 2) {
      std::vector<cd> tmp(in, in + s->N);
      sys_overlap(s, 3, false, tmp.data());
      sys_ham(s, scal, tmp.data(), out);
      sys_overlap(s, 3, true, out);
    } else {
      sys_ham(s, scal, in, out);
      if (mode == 1) sys_overlap(s, 1, false, out);
    }
  });
}

extern "C" int oracle_sys_overlap(void* p, int op, int conj, double* psi) {
  return guard([&] { sys_overlap(static_cast<OracleSystem*>(p), op, conj != 0, reinterpret_cast<cd*>(psi)); });
}

// Exact block-diagonal preconditioner for a rank-3 system: M_c = ss S1xS2 + sh P_c,
// P_c given as (nchan, p1, p2) with Simp pair lists; scalar banded LU without pivoting.
extern "C" int oracle_sys_set_preconditioner(void* p, const double* s1, const double* s2, int band,
                                             const double* pdata, int p1, const int32_t* pairs1, int p2,
                                             const int32_t* pairs2, double sh_re, double sh_im, double ss_re,
                                             double ss_im) {
  return guard([&] {
    auto* s = static_cast<OracleSystem*>(p);
    if (s->ndim != 3) throw std::runtime_error("preconditioner needs rank 3");
    const int nc = s->ext[0], n1 = s->ext[1], n2 = s->ext[2];
    const int n = n1 * n2;
    const int w = band * n2 + band;
    const int ld = 2 * w + 1;
    s->pc_nchan = nc;
    s->pc_n1 = n1;
    s->pc_n2 = n2;
    s->pc_w = w;
    s->pc_lu.assign((size_t)nc * n * ld, 0);
    const cd sh(sh_re, sh_im), ss(ss_re, ss_im);
    const cd* P = reinterpret_cast<const cd*>(pdata);
    std::vector<int> pos1((size_t)n1 * n1, -1), pos2((size_t)n2 * n2, -1);
    for (int q = 0; q < p1; ++q) pos1[(size_t)pairs1[2 * q] * n1 + pairs1[2 * q + 1]] = q;
    for (int q = 0; q < p2; ++q) pos2[(size_t)pairs2[2 * q] * n2 + pairs2[2 * q + 1]] = q;
    auto work = [&](int c) {
      cd* L = s->pc_lu.data() + (size_t)c * n * ld;
      auto E = [&](int i, int j) -> cd& { return L[(size_t)i * ld + w + j - i]; };
      for (int i1 = 0; i1 < n1; ++i1)
        for (int j1 = std::max(0, i1 - band); j1 <= std::min(n1 - 1, i1 + band); ++j1)
          for (int i2 = 0; i2 < n2; ++i2)
            for (int j2 = std::max(0, i2 - band); j2 <= std::min(n2 - 1, i2 + band); ++j2) {
              cd v = ss * (s1[(size_t)i1 * n1 + j1] * s2[(size_t)i2 * n2 + j2]);
              const int q1 = pos1[(size_t)i1 * n1 + j1], q2 = pos2[(size_t)i2 * n2 + j2];
              if (q1 >= 0 && q2 >= 0) v += sh * P[((size_t)c * p1 + q1) * p2 + q2];
              E(i1 * n2 + i2, j1 * n2 + j2) = v;
            }
      for (int k = 0; k < n; ++k) {
        const cd piv = E(k, k);
        if (piv == cd(0)) throw std::runtime_error("zero pivot");
        for (int i = k + 1; i <= std::min(n - 1, k + w); ++i) {
          const cd l = E(i, k) / piv;
          E(i, k) = l;
          if (l == cd(0)) continue;
          for (int j = k + 1; j <= std::min(n - 1, k + w); ++j) E(i, j) -= l * E(k, j);
        }
      }
    };
    unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> pool;
    std::vector<std::string> errs(nt);
    for (unsigned t = 0; t < nt; ++t)
      pool.emplace_back([&, t] {
        try {
          for (int c = (int)t; c < nc; c += (int)nt) work(c);
        } catch (const std::exception& e) {
          errs[t] = e.what();
        }
      });
    for (auto& th : pool) th.join();
    for (auto& e : errs)
      if (!e.empty()) throw std::runtime_error(e);
    s->has_pc = true;
  });
}

static void sys_precond(OracleSystem* s, cd* psi) {
  const int n = s->pc_n1 * s->pc_n2, w = s->pc_w, ld = 2 * w + 1;
  run_parts([&](int part, int nparts) {
  for (int c = part; c < s->pc_nchan; c += nparts) {
    const cd* L = s->pc_lu.data() + (size_t)c * n * ld;
    auto E = [&](int i, int j) { return L[(size_t)i * ld + w + j - i]; };
    cd* x = psi + (size_t)c * n;
    for (int i = 0; i < n; ++i)
      for (int j = std::max(0, i - w); j < i; ++j) x[i] -= E(i, j) * x[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j <= std::min(n - 1, i + w); ++j) x[i] -= E(i, j) * x[j];
      x[i] /= E(i, i);
    }
  }
  });
}

extern "C" int oracle_sys_precond_solve(void* p, double* psi) {
  return guard([&] { sys_precond(static_cast<OracleSystem*>(p), reinterpret_cast<cd*>(psi)); });
}

// dest = (S + hfac H) src  (CayleyPropagator.MultiplySpH, :116-142)
static void multiply_sph(OracleSystem* s, const cd* src, cd* dst, cd hfac) {
  s->t1.assign(src, src + s->N);
  sys_overlap(s, 0, false, s->t1.data());
  sys_ham(s, s->scal.data(), src, dst);
  for (size_t i = 0; i < s->N; ++i) dst[i] = dst[i] * hfac + s->t1[i];
}

static int cayley_cb(void* ctx, const double* in, double* out) {
  auto* s = static_cast<OracleSystem*>(ctx);
  multiply_sph(s, reinterpret_cast<const cd*>(in), reinterpret_cast<cd*>(out), cd(0, 0.5 * s->dt));
  if (s->has_pc) sys_precond(s, reinterpret_cast<cd*>(out));
  return 0;
}

extern "C" int oracle_sys_cayley_step(void* p, double* psi, double dt, const double* scal, int m, double tol,
                                      int dorth, double* out_err, int* out_ops) {
  return guard([&] {
    auto* s = static_cast<OracleSystem*>(p);
    s->dt = dt;
    s->scal.assign(scal, scal + s->terms.size());
    cd* x = reinterpret_cast<cd*>(psi);
    std::vector<cd> b(s->N);
    multiply_sph(s, x, b.data(), cd(0, -0.5 * dt));
    if (s->has_pc) sys_precond(s, b.data());
    GmresResult r = gmres_solve(s->N, m, tol, dorth != 0, &cayley_cb, s, b.data(), x, nullptr);
    if (out_err) *out_err = r.error;
    if (out_ops) *out_ops = r.op_count;
  });
}

static int balanced_cb(void* ctx, const double* in, double* out) {
  auto* s = static_cast<OracleSystem*>(ctx);
  return oracle_sys_multiply(s, 2, s->scal.data(), in, out);
}

extern "C" int oracle_sys_pamp_step(void* p, double* psi, double dt_re, double dt_im, const double* scal, int m,
                                    int order, int dorth, double* err_est, int method) {
  return guard([&] {
    auto* s = static_cast<OracleSystem*>(p);
    s->scal.assign(scal, scal + s->terms.size());
    cd* x = reinterpret_cast<cd*>(psi);
    sys_overlap(s, 2, false, x);
    cd e = pamp_propagate(s->N, m, order, dorth != 0, &balanced_cb, s, x, cd(dt_re, dt_im), nullptr, nullptr, method);
    sys_overlap(s, 3, false, x);
    if (err_est) {
      err_est[0] = e.real();
      err_est[1] = e.imag();
    }
  });
}

extern "C" int oracle_sys_inner_product(void* p, const double* a, const double* b, double* out) {
  return guard([&] {
    auto* s = static_cast<OracleSystem*>(p);
    std::vector<cd> w(reinterpret_cast<const cd*>(b), reinterpret_cast<const cd*>(b) + s->N);
    sys_overlap(s, 0, false, w.data());
    cd r = zdotc(s->N, reinterpret_cast<const cd*>(a), w.data());
    out[0] = r.real();
    out[1] = r.imag();
  });
}

// ------------------------------------------------------------------------------------
// B-spline projection of grid functions (potential generation)
// ------------------------------------------------------------------------------------
struct OracleBSpline {
  int k = 0, N = 0, Q = 0;
  std::vector<double> t, x, w;            // knots; global grid points and weights
  std::vector<int> ivl;                   // interval of each grid point
  std::vector<double> B[3];               // value, first and second derivative, (N, Q)
  std::vector<int> first, count;          // support of each spline on the grid
};

// B(m, i, x) for every i, m = 1..k by the Cox-de Boor recursion; a term whose knot
// difference vanishes is dropped (bspline.cpp:26-76 drops it through its eps guard on B).
static std::vector<std::vector<double>> bspline_levels(const std::vector<double>& t, int k, double x) {
  const int nk = (int)t.size();
  std::vector<std::vector<double>> L(k + 1);
  L[1].assign(nk - 1, 0.0);
  for (int i = 0; i + 1 < nk; ++i) L[1][i] = (t[i] <= x && x < t[i + 1]) ? 1.0 : 0.0;
  for (int m = 2; m <= k; ++m) {
    L[m].assign(nk - m, 0.0);
    for (int i = 0; i + m < nk; ++i) {
      double v = 0;
      const double d0 = t[i + m - 1] - t[i], d1 = t[i + m] - t[i + 1];
      if (d0 > 0) v += (x - t[i]) / d0 * L[m - 1][i];
      if (d1 > 0) v += (t[i + m] - x) / d1 * L[m - 1][i + 1];
      L[m][i] = v;
    }
  }
  return L;
}

static OracleBSpline bspline_setup(int nknots, const double* knots, int k, int nbp, const double* bp, int g,
                                   const double* nodes, const double* weights) {
  OracleBSpline b;
  b.k = k;
  b.t.assign(knots, knots + nknots);
  b.N = nknots - k;
  // global grid: g Gauss points per breakpoint interval, x -> (x - 1)(b - a)/2 + b
  for (int m = 0; m + 1 < nbp; ++m) {
    const double lo = bp[m], hi = bp[m + 1];
    for (int q = 0; q < g; ++q) {
      b.x.push_back((nodes[q] - 1.0) * (hi - lo) / 2.0 + hi);
      b.w.push_back((hi - lo) / 2.0 * weights[q]);
      b.ivl.push_back(m);
    }
  }
  b.Q = (int)b.x.size();
  for (auto& v : b.B) v.assign((size_t)b.N * b.Q, 0.0);
  const auto& t = b.t;
  for (int q = 0; q < b.Q; ++q) {
    const auto L = bspline_levels(t, k, b.x[q]);
    for (int i = 0; i < b.N; ++i) {
      b.B[0][(size_t)i * b.Q + q] = L[k][i];
      // B' = (k-1) [B(k-1,i) / (t[i+k-1]-t[i]) - B(k-1,i+1) / (t[i+k]-t[i+1])]
      double d1 = 0;
      const double a0 = t[i + k - 1] - t[i], a1 = t[i + k] - t[i + 1];
      if (a0 > 0) d1 += (k - 1.0) / a0 * L[k - 1][i];
      if (a1 > 0) d1 -= (k - 1.0) / a1 * L[k - 1][i + 1];
      b.B[1][(size_t)i * b.Q + q] = d1;
      // B'' from differentiating the B' formula once more
      double d2 = 0;
      if (k >= 3) {
        auto dlow = [&](int j) {  // derivative of B(k-1, j)
          double r = 0;
          const double c0 = t[j + k - 2] - t[j], c1 = t[j + k - 1] - t[j + 1];
          if (c0 > 0) r += (k - 2.0) / c0 * L[k - 2][j];
          if (c1 > 0) r -= (k - 2.0) / c1 * L[k - 2][j + 1];
          return r;
        };
        if (a0 > 0) d2 += (k - 1.0) / a0 * dlow(i);
        if (a1 > 0) d2 -= (k - 1.0) / a1 * dlow(i + 1);
      }
      b.B[2][(size_t)i * b.Q + q] = d2;
    }
  }
  // support: the grid points inside [t_i, t_{i+k}] (interior Gauss points only)
  b.first.assign(b.N, 0);
  b.count.assign(b.N, 0);
  for (int i = 0; i < b.N; ++i) {
    int f = -1, c = 0;
    for (int q = 0; q < b.Q; ++q)
      if (b.x[q] > t[i] && b.x[q] < t[i + k]) {
        if (f < 0) f = q;
        ++c;
      }
    b.first[i] = std::max(f, 0);
    b.count[i] = c;
  }
  return b;
}

// BSplineGlobalOverlapIntegral(f, d, row, col): (i, j) = (min, max), factor (-1)^d when
// row > col, zero when j - i >= k, else the sum over the common support of
// w f B_i B_j^(d).
static double overlap_integral(const OracleBSpline& b, const double* f, size_t fstride, int d, int row, int col) {
  int i = row, j = col;
  double factor = 1.0;
  if (i > j) {
    std::swap(i, j);
    factor = (d % 2) ? -1.0 : 1.0;
  }
  if (j - i >= b.k) return 0.0;
  const int lo = std::max(b.first[i], b.first[j]);
  const int hi = std::min(b.first[i] + b.count[i], b.first[j] + b.count[j]);
  double I = 0;
  for (int q = lo; q < hi; ++q)
    I += b.w[q] * f[(size_t)q * fstride] * b.B[0][(size_t)i * b.Q + q] * b.B[d][(size_t)j * b.Q + q];
  return factor * I;
}

// Grid points, weights and value/derivative tables (for checks against other restatements).
extern "C" int oracle_bspline_tables(int nknots, const double* knots, int k, int nbp, const double* bp, int g,
                                     const double* nodes, const double* weights, double* grid, double* gw,
                                     double* b0, double* b1, double* b2) {
  return guard([&] {
    OracleBSpline b = bspline_setup(nknots, knots, k, nbp, bp, g, nodes, weights);
    std::copy(b.x.begin(), b.x.end(), grid);
    std::copy(b.w.begin(), b.w.end(), gw);
    std::copy(b.B[0].begin(), b.B[0].end(), b0);
    std::copy(b.B[1].begin(), b.B[1].end(), b1);
    std::copy(b.B[2].begin(), b.B[2].end(), b2);
  });
}

// One rank of RepresentPotentialInBasisBSpline: src viewed as (pre, Q, post) (MapToRank3),
// dst (pre, pairs, post) = scaling * BSplineGlobalOverlapIntegral(src(pre, :, post), d,
// row, col), scaling 0.5 on the diagonal pairs of "Herm" storage.
extern "C" int oracle_bspline_represent(int nknots, const double* knots, int k, int nbp, const double* bp, int g,
                                        const double* nodes, const double* weights, int d, int npairs,
                                        const int* pairs, int herm, int64_t pre, int64_t post, const double* src,
                                        double* dst) {
  return guard([&] {
    if (d < 0 || d > 2) throw std::runtime_error("invalid derivative");
    OracleBSpline b = bspline_setup(nknots, knots, k, nbp, bp, g, nodes, weights);
    const int Q = b.Q;
    for (int64_t a = 0; a < pre; ++a)
      for (int p = 0; p < npairs; ++p) {
        const int row = pairs[2 * p], col = pairs[2 * p + 1];
        const double sc = (herm && row == col) ? 0.5 : 1.0;
        for (int64_t c = 0; c < post; ++c)
          dst[((size_t)a * npairs + p) * post + c] =
              sc * overlap_integral(b, src + (size_t)a * Q * post + c, (size_t)post, d, row, col);
      }
  });
}

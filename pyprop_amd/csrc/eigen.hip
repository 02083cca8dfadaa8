// Eigenvalue solvers for initial states:
//
//   * pIRAM, the implicitly restarted Arnoldi method of core/krylov/piram/piram/piram.h
//     (Setup 240-312, PerformInitialArnoldiStep 315-345, PerformArnoldiStep 347-396,
//     PerformOrthogonalization 398-444, UpdateEigenvalues 471-518, PerformRestartStep
//     520-610, Solve 640-672): exact shifts at the unwanted Ritz values, which are the
//     ones with the largest real part (CompareComplexLessReal, functors.h:95-102).  The
//     Arnoldi basis and its orthogonalisation live on the device; the m x m Hessenberg
//     work runs on the host, as the reference's LAPACK calls do.
//   * the GMRES shift-invert operator of pyprop/solver/GMRESShiftInvertSolver.py:16-132,
//     x = (H - shift S)^-1 S b, by left-preconditioned GMRES with the block preconditioner
//     set up for H - shift S.
//
// Convergence test: ARPACK's, |y_m| |f| <= tol max(eps^(2/3), |theta|).  The reference
// writes tol max(eps^(2/3), tol |theta|) (piram.h:507), which for its configured tolerance
// 0 (clamped to eps) is never met, so its runs end at the restart limit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"
#include "pp_factored.h"

namespace pp {

using cd = std::complex<double>;

// ------------------------------------------------------------------------------------
// small dense complex linear algebra (host, column-major m x m)
// ------------------------------------------------------------------------------------
struct SmallMat {
  int m = 0;
  std::vector<cd> a;
  explicit SmallMat(int m_ = 0) : m(m_), a((size_t)m_ * m_, 0.0) {}
  cd& operator()(int i, int j) { return a[(size_t)j * m + i]; }
  cd operator()(int i, int j) const { return a[(size_t)j * m + i]; }
  static SmallMat identity(int m) {
    SmallMat I(m);
    for (int i = 0; i < m; ++i) I(i, i) = 1.0;
    return I;
  }
};

// Rotation acting on a pair of rows, [x; y] <- [c s; -conj(s) c] [x; y], chosen to zero y.
struct Rot {
  double c = 1.0;
  cd s = 0.0;
};
static Rot make_rot(cd x, cd y) {
  Rot g;
  const double ax = std::abs(x), ay = std::abs(y);
  if (ay == 0.0) return g;
  if (ax == 0.0) {
    g.c = 0.0;
    g.s = std::conj(y) / ay;
    return g;
  }
  const double r = std::hypot(ax, ay);
  g.c = ax / r;
  g.s = (x / ax) * std::conj(y) / r;
  return g;
}
// rows i, i+1 of A, columns [c0, c1)
static void rot_rows(SmallMat& A, int i, const Rot& g, int c0, int c1) {
  for (int j = c0; j < c1; ++j) {
    const cd x = A(i, j), y = A(i + 1, j);
    A(i, j) = g.c * x + g.s * y;
    A(i + 1, j) = -std::conj(g.s) * x + g.c * y;
  }
}
// columns i, i+1 of A times the adjoint of the rotation, rows [r0, r1)
static void rot_cols_adj(SmallMat& A, int i, const Rot& g, int r0, int r1) {
  for (int k = r0; k < r1; ++k) {
    const cd x = A(k, i), y = A(k, i + 1);
    A(k, i) = g.c * x + std::conj(g.s) * y;
    A(k, i + 1) = -g.s * x + g.c * y;
  }
}

// One explicitly shifted QR step on the whole upper Hessenberg H: H - mu I = Q R,
// H <- R Q + mu I = Q^H H Q, and Qacc <- Qacc Q.
static void shifted_qr_step(SmallMat& H, cd mu, SmallMat& Qacc) {
  const int m = H.m;
  for (int i = 0; i < m; ++i) H(i, i) -= mu;
  std::vector<Rot> rots(std::max(0, m - 1));
  for (int k = 0; k + 1 < m; ++k) {
    rots[k] = make_rot(H(k, k), H(k + 1, k));
    rot_rows(H, k, rots[k], k, m);
    H(k + 1, k) = 0.0;
  }
  for (int k = 0; k + 1 < m; ++k) {
    rot_cols_adj(H, k, rots[k], 0, std::min(k + 2, m));
    rot_cols_adj(Qacc, k, rots[k], 0, Qacc.m);
  }
  for (int i = 0; i < m; ++i) H(i, i) += mu;
}

// Complex Schur form of an upper Hessenberg matrix by shifted QR with deflation:
// H = Z T Z^H on return (H is overwritten by T).  Returns false if it did not converge.
static bool hessenberg_schur(SmallMat& H, SmallMat& Z) {
  const int m = H.m;
  const double eps = std::numeric_limits<double>::epsilon();
  int hi = m - 1, iter = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    while (l > 0) {
      const double scale = std::abs(H(l - 1, l - 1)) + std::abs(H(l, l));
      if (std::abs(H(l, l - 1)) <= eps * (scale > 0 ? scale : 1.0)) break;
      --l;
    }
    if (l > 0) H(l, l - 1) = 0.0;
    if (l == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++total > 100 * m) return false;
    // Wilkinson shift from the trailing 2 x 2 of the active block (every 10th an
    // exceptional shift)
    cd mu;
    if (++iter % 10 == 0) {
      mu = H(hi, hi) + std::abs(H(hi, hi - 1));
    } else {
      const cd a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
      const cd half = 0.5 * (a - d);
      const cd disc = std::sqrt(half * half + b * c);
      const cd m1 = 0.5 * (a + d) + disc, m2 = 0.5 * (a + d) - disc;
      mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
    }
    // QR step on rows/columns l..hi, applied to the full matrix (Schur form) and to Z
    for (int i = l; i <= hi; ++i) H(i, i) -= mu;
    std::vector<Rot> rots(hi - l);
    for (int k = l; k < hi; ++k) {
      rots[k - l] = make_rot(H(k, k), H(k + 1, k));
      rot_rows(H, k, rots[k - l], k, m);
      H(k + 1, k) = 0.0;
    }
    for (int k = l; k < hi; ++k) {
      rot_cols_adj(H, k, rots[k - l], 0, std::min(k + 2, hi + 1));
      rot_cols_adj(Z, k, rots[k - l], 0, m);
    }
    for (int i = l; i <= hi; ++i) H(i, i) += mu;
  }
  return true;
}

// Eigenvalues and unit eigenvectors of an upper Hessenberg matrix (columns of X)
static void hessenberg_eig(const SmallMat& H0, std::vector<cd>& w, SmallMat& X) {
  const int m = H0.m;
  SmallMat T = H0;
  SmallMat Z = SmallMat::identity(m);
  PP_REQUIRE(hessenberg_schur(T, Z), "Hessenberg QR iteration did not converge");
  double tnorm = 0;
  for (auto& z : T.a) tnorm = std::max(tnorm, std::abs(z));
  const double small = std::numeric_limits<double>::epsilon() * (tnorm > 0 ? tnorm : 1.0);
  w.assign(m, 0.0);
  X = SmallMat(m);
  for (int k = 0; k < m; ++k) {
    w[k] = T(k, k);
    std::vector<cd> x(m, 0.0);
    x[k] = 1.0;
    for (int j = k - 1; j >= 0; --j) {
      cd acc = 0.0;
      for (int i = j + 1; i <= k; ++i) acc += T(j, i) * x[i];
      cd den = T(j, j) - w[k];
      if (std::abs(den) < small) den = small;
      x[j] = -acc / den;
    }
    double nrm = 0;
    std::vector<cd> y(m, 0.0);
    for (int i = 0; i < m; ++i) {
      cd acc = 0.0;
      for (int j = 0; j <= k; ++j) acc += Z(i, j) * x[j];
      y[i] = acc;
      nrm += std::norm(acc);
    }
    nrm = std::sqrt(nrm);
    for (int i = 0; i < m; ++i) X(i, k) = y[i] / nrm;
  }
}

// ------------------------------------------------------------------------------------
// pIRAM
// ------------------------------------------------------------------------------------
struct Piram {
  long long n = 0;
  int m = 0, k = 0;  // basis size, wanted eigenvalue count
  double tol = 0;
  int max_restarts = 0, max_orth = 3;
  bool random_start = true;
  DeviceBuffer<double2> V, f, tmp, small, coef;
  Reducer red;
  std::vector<double2> host_small;
  SmallMat H;                   // Hessenberg of the current factorization (row i, column j)
  int cur = -1;                 // columns 0..cur of V are set
  std::vector<cd> ritz;         // wanted Ritz values, ascending real part
  std::vector<double> err_est;  // their error bounds |y_m| |f|
  std::vector<double> conv_est; // error bound - accuracy (<= 0: converged)
  SmallMat Y;                   // Ritz vectors of H (columns, in ritz order for the first k)
  int op_count = 0, restart_count = 0, orth_count = 0;
  bool converged = false;

  void read_small(int count, hipStream_t s) {
    host_small.resize(count);
    PP_HIP(hipMemcpyAsync(host_small.data(), small.ptr, count * sizeof(double2), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
  }
  double2* col(int j) { return V.ptr + (size_t)j * n; }

  void apply(pp_operator_fn op, void* ctx, const double2* in, double2* out, hipStream_t s) {
    ++op_count;
    const int rc = op(ctx, reinterpret_cast<const pp_cplx*>(in), reinterpret_cast<pp_cplx*>(out), (pp_stream)s);
    if (rc != 0) throw Error(PP_ERR_INTERNAL, "pIRAM operator callback failed");
  }

  // f <- f - V_{0..j} h with classical Gram-Schmidt and the reference's re-orthogonalisation
  // rule (repeat while |f| < 0.717 |f before|); returns h, leaves |f| in *fnorm
  std::vector<cd> orthogonalize(int j, double* fnorm, hipStream_t s) {
    const int kk = j + 1;
    block_dot(n, kk, V.ptr, n, f.ptr, small.ptr, red, s);
    block_update(n, kk, V.ptr, n, small.ptr, f.ptr, small.ptr + (kk + 1), red, s);
    read_small(kk + 2, s);
    std::vector<cd> h(kk);
    for (int i = 0; i < kk; ++i) h[i] = cd(host_small[i].x, host_small[i].y);
    double orig = std::sqrt(host_small[kk].x);
    double resid = std::sqrt(host_small[kk + 1].x);
    int count = 0;
    while (resid < 0.717 * orig && count < max_orth) {
      ++count;
      ++orth_count;
      block_dot(n, kk, V.ptr, n, f.ptr, small.ptr, red, s);
      block_update(n, kk, V.ptr, n, small.ptr, f.ptr, small.ptr + (kk + 1), red, s);
      read_small(kk + 2, s);
      for (int i = 0; i < kk; ++i) h[i] += cd(host_small[i].x, host_small[i].y);
      orig = resid;
      resid = std::sqrt(host_small[kk + 1].x);
    }
    *fnorm = resid;
    return h;
  }

  double fnorm(hipStream_t s) {
    vec_norm2(n, f.ptr, small.ptr, red, s);
    read_small(1, s);
    return std::sqrt(host_small[0].x);
  }

  // columns cur+1 .. m-1
  void extend(pp_operator_fn op, void* ctx, hipStream_t s) {
    while (cur < m - 1) {
      const int j = cur + 1;
      const double beta = fnorm(s);
      PP_REQUIRE(beta > 0, "pIRAM: the Krylov space is exhausted (invariant subspace)");
      H(j, j - 1) = beta;
      vec_combine(n, make_double2(1.0 / beta, 0), f.ptr, make_double2(0, 0), nullptr, col(j), s);
      apply(op, ctx, col(j), f.ptr, s);
      double fn = 0;
      const std::vector<cd> h = orthogonalize(j, &fn, s);
      for (int i = 0; i <= j; ++i) H(i, j) = h[i];
      cur = j;
    }
  }

  void start(pp_operator_fn op, void* ctx, const double2* v0, hipStream_t s) {
    if (v0) {
      vec_copy(n, v0, f.ptr, s);
    } else {
      // uniform [0, 1) real start (SetupRandomResidual, piram.h:624-636), fixed seed
      std::vector<double2> r((size_t)n);
      unsigned long long st = 0x9E3779B97F4A7C15ull;
      for (auto& z : r) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z = make_double2((double)(st >> 11) * (1.0 / 9007199254740992.0), 0.0);
      }
      tmp.upload(r.data(), r.size(), s);
      vec_copy(n, tmp.ptr, f.ptr, s);
    }
    const double nr = fnorm(s);
    PP_REQUIRE(nr > 0, "pIRAM start vector is zero");
    vec_combine(n, make_double2(1.0 / nr, 0), f.ptr, make_double2(0, 0), nullptr, col(0), s);
    apply(op, ctx, col(0), f.ptr, s);
    double fn = 0;
    const std::vector<cd> h = orthogonalize(0, &fn, s);
    H = SmallMat(m);
    H(0, 0) = h[0];
    cur = 0;
  }

  // Ritz values/vectors, error bounds; returns the unwanted Ritz values as shifts, the one
  // with the smallest error bound first (UpdateEigenvalues, piram.h:471-518)
  std::vector<cd> update(hipStream_t s) {
    const double fn = fnorm(s);
    std::vector<cd> w;
    SmallMat X;
    hessenberg_eig(H, w, X);
    std::vector<int> order(m);
    for (int i = 0; i < m; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a].real() < w[b].real(); });
    const double eps = std::numeric_limits<double>::epsilon();
    const double eps23 = std::pow(eps, 2.0 / 3.0);
    ritz.assign(k, 0.0);
    err_est.assign(k, 0.0);
    conv_est.assign(k, 0.0);
    Y = SmallMat(m);
    converged = true;
    for (int i = 0; i < m; ++i)
      for (int r = 0; r < m; ++r) Y(r, i) = X(r, order[i]);
    for (int i = 0; i < k; ++i) {
      const int o = order[i];
      ritz[i] = w[o];
      err_est[i] = std::abs(X(m - 1, o)) * fn;
      conv_est[i] = err_est[i] - tol * std::max(eps23, std::abs(w[o]));
      converged = converged && conv_est[i] <= 0.0;
    }
    std::vector<std::pair<double, cd>> sh;
    for (int i = k; i < m; ++i) sh.push_back({std::abs(X(m - 1, order[i])) * fn, w[order[i]]});
    std::stable_sort(sh.begin(), sh.end(), [](const std::pair<double, cd>& a, const std::pair<double, cd>& b) {
      return a.first < b.first;
    });
    std::vector<cd> shifts;
    for (auto& p : sh) shifts.push_back(p.second);
    return shifts;
  }

  // PerformRestartStep (piram.h:520-610): p = m - k shifted QR steps, then compress the
  // factorization to k columns
  void restart(const std::vector<cd>& shifts, hipStream_t s) {
    ++restart_count;
    SmallMat Q = SmallMat::identity(m);
    for (const cd& mu : shifts) shifted_qr_step(H, mu, Q);
    // new columns 0..k: (V Q)_i; computed into tmp rows then copied back
    if (tmp.count < (size_t)(k + 1) * n) tmp.resize((size_t)(k + 1) * n);
    std::vector<double2> c((size_t)m);
    for (int i = 0; i <= k; ++i) {
      for (int j = 0; j < m; ++j) c[j] = make_double2(Q(j, i).real(), Q(j, i).imag());
      coef.upload(c.data(), c.size(), s);
      vec_lincomb(n, m, V.ptr, n, coef.ptr, nullptr, tmp.ptr + (size_t)i * n, s);
    }
    // f <- (V Q)_k H(k, k-1) + f Q(m-1, k-1)
    const cd beta = H(k, k - 1), sigma = Q(m - 1, k - 1);
    vec_combine(n, make_double2(beta.real(), beta.imag()), tmp.ptr + (size_t)k * n,
                make_double2(sigma.real(), sigma.imag()), f.ptr, f.ptr, s);
    PP_HIP(hipMemcpyAsync(V.ptr, tmp.ptr, (size_t)k * n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    SmallMat Hk(m);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) Hk(i, j) = H(i, j);
    H = Hk;
    cur = k - 1;
  }

  void solve(pp_operator_fn op, void* ctx, const double2* v0, hipStream_t s) {
    op_count = restart_count = orth_count = 0;
    start(op, ctx, v0, s);
    extend(op, ctx, s);
    while (true) {
      const std::vector<cd> shifts = update(s);
      if (converged || restart_count >= max_restarts) break;
      restart(shifts, s);
      extend(op, ctx, s);
    }
  }

  void eigenvector(int i, double2* out, hipStream_t s) {
    PP_REQUIRE(i >= 0 && i < k && !ritz.empty(), "eigenvector index out of range (solve first)");
    std::vector<double2> c((size_t)m);
    for (int j = 0; j < m; ++j) c[j] = make_double2(Y(j, i).real(), Y(j, i).imag());
    coef.upload(c.data(), c.size(), s);
    vec_lincomb(n, m, V.ptr, n, coef.ptr, nullptr, out, s);
  }
};

// ------------------------------------------------------------------------------------
// GMRES shift-invert: x = (H - shift S)^-1 S b
// ------------------------------------------------------------------------------------
struct ShiftInvert {
  Hamiltonian* ham = nullptr;
  const FHam* fham = nullptr;
  const Precond* pc = nullptr;
  double shift = 0;
  std::vector<double> scalings;
  Gmres gmres;
  DeviceBuffer<double2> tmp, rhs;
  int solves = 0, op_total = 0;
  bool negate = false;  // the operator form (pp_shift_invert_apply) returns -(H - shift S)^-1 S
  ShiftInvert(Hamiltonian* h, const Precond* p, int m, double tol, bool dorth, double sh)
      : ham(h), pc(p), shift(sh), scalings(h->terms.size(), 1.0), gmres(h->n, m, tol, dorth) {
    tmp.resize(h->n);
    rhs.resize(h->n);
  }
  // __callback (GMRESShiftInvertSolver.py:60-97): out = M^-1 (H in - shift S in)
  static int op_cb(void* ctx, const pp_cplx* in_, pp_cplx* out_, pp_stream s_) {
    auto* si = static_cast<ShiftInvert*>(ctx);
    try {
      const double2* in = reinterpret_cast<const double2*>(in_);
      double2* out = reinterpret_cast<double2*>(out_);
      hipStream_t s = (hipStream_t)s_;
      const long long n = si->ham->n;
      vec_copy(n, in, si->tmp.ptr, s);
      si->ham->repr->overlap(PP_OVERLAP_MULTIPLY, false, si->tmp.ptr, s);
      vec_combine(n, make_double2(-si->shift, 0), si->tmp.ptr, make_double2(0, 0), nullptr, si->tmp.ptr, s);
      if (si->fham) fham_apply64(*si->fham, si->scalings.data(), in, out, si->tmp.ptr, make_double2(1.0, 0), s);
      else si->ham->apply(si->scalings.data(), in, out, si->tmp.ptr, make_double2(1.0, 0), s);
      if (si->pc) precond_solve(*si->pc, out, s);
      return 0;
    } catch (const Error& e) {
      set_last_error(e.what());
      return e.code;
    }
  }
  // InverseIterations (GMRESShiftInvertSolver.py:100-126)
  void solve(const double2* b, double2* x, hipStream_t s) {
    const long long n = ham->n;
    vec_copy(n, b, rhs.ptr, s);
    ham->repr->overlap(PP_OVERLAP_MULTIPLY, false, rhs.ptr, s);
    if (pc) precond_solve(*pc, rhs.ptr, s);
    gmres.solve(&ShiftInvert::op_cb, this, rhs.ptr, x, s);
    ++solves;
    op_total += gmres.op_count;
  }
};

// Hamiltonian as an operator (MultiplyHamiltonianNoOverlap / MultiplyHamiltonian /
// MultiplyHamiltonianBalancedOverlap, BasisPropagator.py:116-159)
struct HamOp {
  Hamiltonian* ham = nullptr;
  std::vector<double> scalings;
  int mode = 0;
  static int op_cb(void* ctx, const pp_cplx* in_, pp_cplx* out_, pp_stream s_) {
    auto* h = static_cast<HamOp*>(ctx);
    try {
      const double2* in = reinterpret_cast<const double2*>(in_);
      double2* out = reinterpret_cast<double2*>(out_);
      hipStream_t s = (hipStream_t)s_;
      if (h->mode == 2) {
        h->ham->multiply_balanced(h->scalings.data(), in, out, s);
      } else {
        h->ham->apply(h->scalings.data(), in, out, nullptr, make_double2(1.0, 0), s);
        if (h->mode == 1) h->ham->repr->overlap(PP_OVERLAP_SOLVE, false, out, s);
      }
      return 0;
    } catch (const Error& e) {
      set_last_error(e.what());
      return e.code;
    }
  }
};

}  // namespace pp

using namespace pp;

struct pp_piram {
  Piram p;
};
struct pp_shift_invert {
  ShiftInvert si;
  pp_shift_invert(Hamiltonian* h, const Precond* p, int m, double tol, bool dorth, double sh)
      : si(h, p, m, tol, dorth, sh) {}
};
struct pp_ham_op {
  HamOp op;
};

extern "C" int pp_piram_create(int64_t n, int basis_size, int eigenvalue_count, double tolerance, int max_restarts,
                               int max_orthogonalization, pp_piram** out) {
  return guarded([&] {
    PP_REQUIRE(out && n > 0, "invalid argument");
    PP_REQUIRE(eigenvalue_count >= 1 && 2 * eigenvalue_count <= basis_size && basis_size <= 128,
               "basis size must be at least twice the eigenvalue count (piram.h:269-272) and at most 128");
    PP_REQUIRE(max_restarts >= 0 && max_orthogonalization >= 0, "invalid restart or orthogonalization limit");
    auto* p = new pp_piram();
    Piram& P = p->p;
    P.n = n;
    P.m = basis_size;
    P.k = eigenvalue_count;
    P.tol = std::max(tolerance, std::numeric_limits<double>::epsilon());
    P.max_restarts = max_restarts;
    P.max_orth = max_orthogonalization;
    try {
      P.V.resize((size_t)basis_size * n);
      P.f.resize(n);
      P.tmp.resize(n);
      P.small.resize(2 * (basis_size + 4));
      P.coef.resize(basis_size);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

extern "C" int pp_piram_destroy(pp_piram* p) {
  return guarded([&] { delete p; });
}

extern "C" int pp_piram_solve(pp_piram* p, pp_operator_fn op, void* ctx, const pp_cplx* start, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && op, "null argument");
    p->p.solve(op, ctx, reinterpret_cast<const double2*>(start), (hipStream_t)stream);
  });
}

extern "C" int pp_piram_eigenvalues(const pp_piram* p, pp_cplx* values, double* error_estimates,
                                    double* convergence_estimates, int* converged_count) {
  return guarded([&] {
    PP_REQUIRE(p && !p->p.ritz.empty(), "pIRAM has not been solved");
    const Piram& P = p->p;
    int conv = 0;
    for (int i = 0; i < P.k; ++i) {
      if (values) values[i] = pp_cplx{P.ritz[i].real(), P.ritz[i].imag()};
      if (error_estimates) error_estimates[i] = P.err_est[i];
      if (convergence_estimates) convergence_estimates[i] = P.conv_est[i];
      if (P.conv_est[i] <= 0.0) ++conv;
    }
    if (converged_count) *converged_count = conv;
  });
}

extern "C" int pp_piram_eigenvector(pp_piram* p, int index, pp_cplx* out, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && out, "null argument");
    p->p.eigenvector(index, reinterpret_cast<double2*>(out), (hipStream_t)stream);
  });
}

extern "C" int pp_piram_stats(const pp_piram* p, int* operator_count, int* restart_count,
                              int* orthogonalization_count) {
  return guarded([&] {
    PP_REQUIRE(p, "null argument");
    if (operator_count) *operator_count = p->p.op_count;
    if (restart_count) *restart_count = p->p.restart_count;
    if (orthogonalization_count) *orthogonalization_count = p->p.orth_count;
  });
}

extern "C" int pp_shift_invert_create(pp_hamiltonian* ham, const pp_precond* pc, int basis_size, double tolerance,
                                      int double_orthogonalization, double shift, pp_shift_invert** out) {
  return guarded([&] {
    PP_REQUIRE(out && ham, "null argument");
    *out = new pp_shift_invert(&ham->h, pc ? &pc->p : nullptr, basis_size, tolerance,
                               double_orthogonalization != 0, shift);
  });
}

extern "C" int pp_shift_invert_destroy(pp_shift_invert* si) {
  return guarded([&] { delete si; });
}

extern "C" int pp_shift_invert_set_scalings(pp_shift_invert* si, const double* scalings) {
  return guarded([&] {
    PP_REQUIRE(si && scalings, "null argument");
    si->si.scalings.assign(scalings, scalings + si->si.ham->terms.size());
  });
}

extern "C" int pp_shift_invert_set_factored(pp_shift_invert* si, const pp_fham* f) {
  return guarded([&] {
    PP_REQUIRE(si, "null argument");
    si->si.fham = f ? &f->f : nullptr;
  });
}

extern "C" int pp_shift_invert_solve(pp_shift_invert* si, const pp_cplx* src, pp_cplx* dst, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(si && src && dst, "null argument");
    si->si.solve(reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst), (hipStream_t)stream);
  });
}

extern "C" int pp_shift_invert_apply(void* ctx, const pp_cplx* in, pp_cplx* out, pp_stream stream) {
  auto* si = static_cast<pp_shift_invert*>(ctx);
  return guarded([&] {
    PP_REQUIRE(si && in && out, "null argument");
    double2* o = reinterpret_cast<double2*>(out);
    si->si.solve(reinterpret_cast<const double2*>(in), o, (hipStream_t)stream);
    if (si->si.negate)
      vec_combine(si->si.ham->n, make_double2(-1.0, 0.0), o, make_double2(0.0, 0.0), nullptr, o, (hipStream_t)stream);
  });
}

extern "C" int pp_shift_invert_set_negate(pp_shift_invert* si, int negate) {
  return guarded([&] {
    PP_REQUIRE(si, "null argument");
    si->si.negate = negate != 0;
  });
}

extern "C" int pp_shift_invert_stats(const pp_shift_invert* si, int* solves, int* operator_count,
                                     double* last_error) {
  return guarded([&] {
    PP_REQUIRE(si, "null argument");
    if (solves) *solves = si->si.solves;
    if (operator_count) *operator_count = si->si.op_total;
    if (last_error) *last_error = si->si.gmres.error;
  });
}

extern "C" int pp_ham_op_create(pp_hamiltonian* ham, const double* scalings, int mode, pp_ham_op** out) {
  return guarded([&] {
    PP_REQUIRE(out && ham && mode >= 0 && mode <= 2, "invalid argument");
    auto* h = new pp_ham_op();
    h->op.ham = &ham->h;
    h->op.mode = mode;
    if (scalings) h->op.scalings.assign(scalings, scalings + h->op.ham->terms.size());
    else h->op.scalings.assign(h->op.ham->terms.size(), 1.0);
    *out = h;
  });
}

extern "C" int pp_ham_op_destroy(pp_ham_op* h) {
  return guarded([&] { delete h; });
}

extern "C" int pp_ham_op_apply(void* ctx, const pp_cplx* in, pp_cplx* out, pp_stream stream) {
  return HamOp::op_cb(&static_cast<pp_ham_op*>(ctx)->op, in, out, stream);
}

// Solver-level internal types: representation, Hamiltonian, GMRES, pAMP, Cayley,
// block-diagonal preconditioner.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"

namespace pp {

void block_dot_if(long long n, int m, const double2* V, long long ld, const double2* w, double2* out_dev,
                  Reducer& red, const int* cond, hipStream_t s);
void block_update_if(long long n, int m, const double2* V, long long ld, const double2* h_dev, double2* w,
                     double2* norm2_dev, Reducer& red, const int* cond, hipStream_t s);

struct Repr {
  int ndim = 0;
  std::vector<int> extents;
  std::vector<const OverlapFactor*> ov;
  long long size() const {
    long long s = 1;
    for (int e : extents) s *= e;
    return s;
  }
  // work: out-of-place buffer of size() elements for the multiply ops (NULL: each factor's own)
  void overlap(int op, bool conj, double2* psi, hipStream_t s, double2* work = nullptr) const;
};

struct Hamiltonian {
  const Repr* repr = nullptr;
  long long n = 0;
  std::vector<const TPTerm*> terms;
  bool fused = false;
  TPPlanDevice fused_plan;
  DeviceBuffer<double2> work;
  // dst = alpha * sum_p scalings[p] V_p src + addend
  void apply(const double* scalings, const double2* src, double2* dst, const double2* addend, double2 alpha,
             hipStream_t s);
  void apply_f32(const double* scalings, const float2* src, float2* dst, const float2* addend, float2 alpha,
                 hipStream_t s);
  void multiply_balanced(const double* scalings, const double2* src, double2* dst, hipStream_t s);
  void ensure_work();
};

struct Gmres {
  long long n;
  int m;
  double tol;
  bool dorth;
  int max_orth = 5;
  DeviceBuffer<double2> V, r, small, coef;
  Reducer red;
  std::vector<std::complex<double>> H, y;
  std::vector<double2> host_small;
  std::vector<double> err_list;
  double error = 0;
  int op_count = 0, orth_count = 0, steps = 0;
  Gmres(long long n, int m, double tol, bool dorth);
  void read_small(int count, hipStream_t s);
  void solve(pp_operator_fn op, void* ctx, const double2* rhs, double2* x, hipStream_t s);
};

struct Pamp {
  long long n;
  int m;
  int order;
  bool dorth;
  int max_orth = 3;
  double tol = std::sqrt(std::numeric_limits<double>::epsilon());
  DeviceBuffer<double2> V, r, H, sc, hbuf, expcol, work;
  DeviceBuffer<int> flags, piv;
  Reducer red;
  int op_count = 0;
  int method = 0;  // 0: Pade exponential on the device, 1: eigenvector exponential on the host
  Pamp(long long n, int m, int order, bool dorth);
  void propagate(pp_operator_fn op, void* ctx, double2* v, double2 dt, hipStream_t s);
};

struct Precond {
  int nchan = 0, n1 = 0, n2 = 0, b = 0;
  // per channel c, block row I (n2 x n2 row-major blocks):
  DeviceBuffer<double2> linv;  // [c][I]      L_II^-1 (unit lower)
  DeviceBuffer<double2> lp;    // [c][I][row][q*n2 + j] = (L_II^-1 L_{I,I-b+q})(row, j)
  DeviceBuffer<double2> uinv;  // [c][I]      U_II^-1
  DeviceBuffer<double2> up;    // [c][I][row][q*n2 + j] = (U_II^-1 U_{I,I+1+q})(row, j)
};
void precond_solve(const Precond& pc, double2* psi, hipStream_t s);

struct Kron;  // separable (Kronecker) preconditioner, mixed.hip
struct FHam;  // factored Hamiltonian, factored.hip
// dst = addend + alpha * H src through the factored operator (FP64)
void fham_apply64(const FHam& f, const double* scalings, const double2* src, double2* dst, const double2* addend,
                  double2 alpha, hipStream_t s);
// psi <- K^-1 psi with K_c = S x S + a (h_l1 x S + S x h_l2) per channel (FP64)
void kron_solve(const Kron& k, double2 a, double2* psi, hipStream_t s);

struct Cayley {
  Hamiltonian* ham = nullptr;
  const Precond* pc = nullptr;
  const Kron* kron = nullptr;  // used when pc is null
  const FHam* fham = nullptr;  // factored multiply in place of the tensor potentials
  long long n = 0;
  Gmres gmres;
  DeviceBuffer<double2> tmp_s, rhs;
  double dt = 0;
  std::vector<double> scalings;
  Cayley(Hamiltonian* h, const Precond* p, int m, double tol, bool dorth)
      : ham(h), pc(p), n(h->n), gmres(h->n, m, tol, dorth) {
    tmp_s.resize(n);
    rhs.resize(n);
  }
  static int operator_cb(void* ctx, const pp_cplx* in, pp_cplx* out, pp_stream s);
  void step(double2* psi, double dt, const double* scalings, hipStream_t s);
};

void pamp_step_balanced(Pamp& p, Hamiltonian& ham, double2* psi, double2 dt, const double* scalings, hipStream_t s);

}  // namespace pp

struct pp_repr {
  pp::Repr r;
};
struct pp_hamiltonian {
  pp::Hamiltonian h;
};
struct pp_gmres {
  pp::Gmres g;
  explicit pp_gmres(long long n, int m, double tol, bool d) : g(n, m, tol, d) {}
};
struct pp_pamp {
  pp::Pamp p;
  explicit pp_pamp(long long n, int m, int o, bool d) : p(n, m, o, d) {}
};
struct pp_precond {
  pp::Precond p;
};
struct pp_cayley {
  pp::Cayley c;
  pp_cayley(pp::Hamiltonian* h, const pp::Precond* p, int m, double tol, bool d) : c(h, p, m, tol, d) {}
};

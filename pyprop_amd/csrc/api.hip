// C ABI entry points for representations, Hamiltonians and the Krylov propagators.
#include <hip/hip_runtime.h>

#include <string>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"

namespace pp {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace pp

using namespace pp;

extern "C" const char* pp_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* pp_version(void) { return "pyprop_amd 0.1 (gfx950)"; }

extern "C" int pp_stream_synchronize(pp_stream stream) {
  return guarded([&] { PP_HIP(hipStreamSynchronize((hipStream_t)stream)); });
}

// ---------------------------------------------------------------- representation
extern "C" int pp_repr_create(int ndim, const int* extents, pp_overlap* const* overlaps, pp_repr** out) {
  return guarded([&] {
    PP_REQUIRE(out && extents && ndim >= 1 && ndim <= kMaxRank, "invalid representation");
    auto* r = new pp_repr();
    r->r.ndim = ndim;
    r->r.extents.assign(extents, extents + ndim);
    r->r.ov.assign(ndim, nullptr);
    for (int i = 0; i < ndim; ++i) {
      if (overlaps && overlaps[i]) {
        if (overlaps[i]->f.n != extents[i]) {
          delete r;
          throw Error(PP_ERR_INVALID, "overlap size does not match extent of rank " + std::to_string(i));
        }
        r->r.ov[i] = &overlaps[i]->f;
      }
    }
    *out = r;
  });
}

extern "C" int pp_repr_destroy(pp_repr* r) {
  return guarded([&] { delete r; });
}

extern "C" int pp_repr_overlap(const pp_repr* repr, int op, int conj, pp_cplx* psi, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(repr && psi, "NULL argument");
    PP_REQUIRE(op >= 0 && op <= 3, "unknown overlap op");
    repr->r.overlap(op, conj != 0, reinterpret_cast<double2*>(psi), (hipStream_t)stream);
  });
}

extern "C" int pp_repr_inner_product(const pp_repr* repr, const pp_cplx* a, const pp_cplx* b, pp_cplx* work,
                                     pp_cplx* result, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(repr && a && b && work && result, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const long long n = repr->r.size();
    double2* w = reinterpret_cast<double2*>(work);
    vec_copy(n, reinterpret_cast<const double2*>(b), w, s);
    repr->r.overlap(PP_OVERLAP_MULTIPLY, false, w, s);
    Reducer red;
    red.ensure(reduction_blocks(n), 2);
    vec_dot(n, reinterpret_cast<const double2*>(a), w, red.result.ptr, red, s);
    double2 h;
    PP_HIP(hipMemcpyAsync(&h, red.result.ptr, sizeof(h), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    result->re = h.x;
    result->im = h.y;
  });
}

// ---------------------------------------------------------------- Hamiltonian
extern "C" int pp_ham_create(const pp_repr* repr, int count, pp_tensor_potential* const* potentials,
                             pp_hamiltonian** out) {
  return guarded([&] {
    PP_REQUIRE(repr && out && potentials && count >= 1 && count <= kMaxTerms, "1..8 potentials per Hamiltonian");
    auto* h = new pp_hamiltonian();
    try {
      h->h.repr = &repr->r;
      h->h.n = repr->r.size();
      for (int i = 0; i < count; ++i) {
        PP_REQUIRE(potentials[i], "NULL potential");
        const TPGeometry& g = potentials[i]->term.geom;
        PP_REQUIRE(g.rank == repr->r.ndim, "potential rank differs from the representation");
        for (int r = 0; r < g.rank; ++r)
          PP_REQUIRE(g.src_ext[r] == repr->r.extents[r], "potential shape differs from the representation");
        h->h.terms.push_back(&potentials[i]->term);
      }
      h->h.fused = terms_fusable(h->h.terms);
      if (h->h.fused) {
        TPPlanData pd = build_plan_host(h->h.terms);
        upload_plan(pd, h->h.fused_plan, nullptr);
        PP_HIP(hipDeviceSynchronize());
      }
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

extern "C" int pp_ham_destroy(pp_hamiltonian* h) {
  return guarded([&] { delete h; });
}

extern "C" int pp_ham_multiply_nooverlap(pp_hamiltonian* ham, const double* scalings, const pp_cplx* src,
                                         pp_cplx* dest, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(ham && src && dest, "NULL argument");
    ham->h.apply(scalings, reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dest), nullptr,
                 make_double2(1.0, 0.0), (hipStream_t)stream);
  });
}

extern "C" int pp_ham_multiply(pp_hamiltonian* ham, const double* scalings, const pp_cplx* src, pp_cplx* dest,
                               pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(ham && src && dest, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    ham->h.apply(scalings, reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dest), nullptr,
                 make_double2(1.0, 0.0), s);
    ham->h.repr->overlap(PP_OVERLAP_SOLVE, false, reinterpret_cast<double2*>(dest), s);
  });
}

extern "C" int pp_ham_multiply_balanced(pp_hamiltonian* ham, const double* scalings, const pp_cplx* src,
                                        pp_cplx* dest, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(ham && src && dest, "NULL argument");
    ham->h.multiply_balanced(scalings, reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dest),
                             (hipStream_t)stream);
  });
}

// ---------------------------------------------------------------- GMRES
extern "C" int pp_gmres_create(int64_t n, int basis_size, double tolerance, int double_orthogonalization,
                               pp_gmres** out) {
  return guarded([&] {
    PP_REQUIRE(out && n > 0, "invalid GMRES arguments");
    *out = new pp_gmres(n, basis_size, tolerance, double_orthogonalization != 0);
  });
}

extern "C" int pp_gmres_destroy(pp_gmres* g) {
  return guarded([&] { delete g; });
}

extern "C" int pp_gmres_solve(pp_gmres* g, pp_operator_fn op, void* ctx, const pp_cplx* rhs, pp_cplx* x,
                              pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(g && op && rhs && x, "NULL argument");
    g->g.solve(op, ctx, reinterpret_cast<const double2*>(rhs), reinterpret_cast<double2*>(x), (hipStream_t)stream);
  });
}

extern "C" int pp_gmres_stats(const pp_gmres* g, double* error_estimate, int* operator_count,
                              int* orthogonalization_count) {
  return guarded([&] {
    PP_REQUIRE(g, "NULL argument");
    if (error_estimate) *error_estimate = g->g.error;
    if (operator_count) *operator_count = g->g.op_count;
    if (orthogonalization_count) *orthogonalization_count = g->g.orth_count;
  });
}

extern "C" int pp_gmres_error_list(const pp_gmres* g, double* out) {
  return guarded([&] {
    PP_REQUIRE(g && out, "NULL argument");
    std::copy(g->g.err_list.begin(), g->g.err_list.end(), out);
  });
}

// ---------------------------------------------------------------- pAMP
extern "C" int pp_pamp_create(int64_t n, int basis_size, int pade_order, int double_orthogonalization,
                              pp_pamp** out) {
  return guarded([&] {
    PP_REQUIRE(out && n > 0, "invalid pAMP arguments");
    *out = new pp_pamp(n, basis_size, pade_order, double_orthogonalization != 0);
  });
}

extern "C" int pp_pamp_destroy(pp_pamp* p) {
  return guarded([&] { delete p; });
}

extern "C" int pp_pamp_set_exponentiation(pp_pamp* p, int method) {
  return guarded([&] {
    PP_REQUIRE(p && (method == 0 || method == 1), "exponentiation method must be 0 (Pade) or 1 (eigenvector)");
    p->p.method = method;
  });
}

extern "C" int pp_pamp_propagate(pp_pamp* p, pp_operator_fn op, void* ctx, pp_cplx* v, pp_cplx dt,
                                 pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && op && v, "NULL argument");
    p->p.propagate(op, ctx, reinterpret_cast<double2*>(v), make_double2(dt.re, dt.im), (hipStream_t)stream);
  });
}

extern "C" int pp_pamp_stats(const pp_pamp* p, pp_cplx* error_estimate, int* breakdown, int* operator_count) {
  return guarded([&] {
    PP_REQUIRE(p, "NULL argument");
    double2 sc[3];
    int fl[2];
    PP_HIP(hipDeviceSynchronize());
    PP_HIP(hipMemcpy(sc, p->p.sc.ptr, sizeof(sc), hipMemcpyDeviceToHost));
    PP_HIP(hipMemcpy(fl, p->p.flags.ptr, sizeof(fl), hipMemcpyDeviceToHost));
    if (error_estimate) {
      error_estimate->re = sc[2].x;
      error_estimate->im = sc[2].y;
    }
    if (breakdown) *breakdown = fl[0];
    if (operator_count) *operator_count = p->p.op_count;
  });
}

extern "C" int pp_pamp_step_balanced(pp_pamp* p, pp_hamiltonian* ham, pp_cplx* psi, pp_cplx dt,
                                     const double* scalings, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(p && ham && psi, "NULL argument");
    PP_REQUIRE(p->p.n == ham->h.n, "pAMP size differs from the Hamiltonian");
    pamp_step_balanced(p->p, ham->h, reinterpret_cast<double2*>(psi), make_double2(dt.re, dt.im), scalings,
                       (hipStream_t)stream);
  });
}

// ---------------------------------------------------------------- Cayley
extern "C" int pp_cayley_create(pp_hamiltonian* ham, const pp_precond* pc, int basis_size, double tolerance,
                                int double_orthogonalization, pp_cayley** out) {
  return guarded([&] {
    PP_REQUIRE(ham && out, "NULL argument");
    if (pc) {
      const Repr& r = *ham->h.repr;
      PP_REQUIRE(r.ndim == 3 && r.extents[0] == pc->p.nchan && r.extents[1] == pc->p.n1 && r.extents[2] == pc->p.n2,
                 "preconditioner shape differs from the wavefunction");
    }
    *out = new pp_cayley(&ham->h, pc ? &pc->p : nullptr, basis_size, tolerance, double_orthogonalization != 0);
  });
}

extern "C" int pp_cayley_destroy(pp_cayley* c) {
  return guarded([&] { delete c; });
}

extern "C" int pp_cayley_step(pp_cayley* c, pp_cplx* psi, double dt, const double* scalings, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(c && psi && scalings, "NULL argument");
    c->c.step(reinterpret_cast<double2*>(psi), dt, scalings, (hipStream_t)stream);
  });
}

extern "C" int pp_cayley_stats(const pp_cayley* c, double* error_estimate, int* operator_count) {
  return guarded([&] {
    PP_REQUIRE(c, "NULL argument");
    if (error_estimate) *error_estimate = c->c.gmres.error;
    if (operator_count) *operator_count = c->c.gmres.op_count;
  });
}

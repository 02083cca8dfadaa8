/*
 * pyprop_amd: MI355X (gfx950) implementation of PyProp's basis-propagation hot path.
 *
 * C ABI.  Plain pointers and sizes only; complex numbers are pp_cplx (two doubles,
 * layout-compatible with C99 double complex and std::complex<double>).  Arrays are
 * C-ordered with extent 0 the slowest rank, exactly as the reference passes blitz
 * arrays to its generated Fortran kernels.  Device pointers are HIP device memory on
 * the current device; every call is stream-ordered on the hipStream_t it receives
 * (NULL = legacy default stream).  Functions return 0 (PP_OK) or a negative error
 * code; pp_last_error() gives the message.  Buffers belong to the caller unless a
 * function says otherwise; library objects (plans, solvers) own their internal
 * device memory until the matching *_destroy.
 *
 * Each declaration cites the reference interface it replaces (paths relative to the
 * PyProp source tree).
 */
#ifndef PYPROP_AMD_H
#define PYPROP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pp_cplx {
  double re;
  double im;
} pp_cplx;

typedef struct pp_cplxf {
  float re;
  float im;
} pp_cplxf;

typedef void* pp_stream; /* hipStream_t */

enum {
  PP_OK = 0,
  PP_ERR_INVALID = -1,
  PP_ERR_HIP = -2,
  PP_ERR_ALLOC = -3,
  PP_ERR_INTERNAL = -4,
  PP_ERR_NOT_CONVERGED = -5
};

/* Storage ids of one tensor-potential rank
 * (core/tensorpotential/tensorpotentialmultiply_generator.py:1336-1346). */
enum {
  PP_STORAGE_SIMP = 0,   /* pair list                            generator.py:302-363 */
  PP_STORAGE_HERM = 1,   /* pair list + conjugated transpose     generator.py:370-398 */
  PP_STORAGE_DIAG = 2,   /* row == col == potential index        generator.py:403-432 */
  PP_STORAGE_IDENT = 3,  /* potential index 0, row == col        generator.py:436-462 */
  PP_STORAGE_BAND = 4,   /* zhbmv 'L' band, lda = k              generator.py:467-618 */
  PP_STORAGE_BANDNH = 5, /* zgbmv band, lda = 2b+1               generator.py:622-762 */
  PP_STORAGE_DENSE = 6   /* row-major N x N                      generator.py:1227-1330 */
};

/* Library-level */
const char* pp_last_error(void);
const char* pp_version(void);
/* Synchronize a stream (convenience for FFI callers without HIP bindings). */
int pp_stream_synchronize(pp_stream stream);

/* ------------------------------------------------------------------------------
 * Tensor potential multiply
 * ------------------------------------------------------------------------------ */

/* One-shot drop-in for the generated kernels
 *   tensorpotentialmultiply_<ids>_(potential, potentialExtent0.., scaling, source, ...,
 *                                  dest, ..., pair0, pair0Extent0, pair0Extent1, ...)
 * (parameter meaning: generator.py:1373-1385 and 1428-1451; C++ wrapper
 * TensorPotentialMultiply_<ids>_Wrapper, generator.py:1486-1493).
 *   dest += V * scaling * source  (dest accumulates; the caller zeroes it)
 * storage_ids[r]: PP_STORAGE_*; pairs[r]: host int32 [pair_counts[r]][2] (row, col)
 * for SIMP/HERM ranks, ignored otherwise.  potential, source, dest: device.
 * Builds an internal plan, runs it on `stream`, synchronizes the stream and frees the
 * plan: for repeated products use pp_tp_create/pp_tp_multiply. */
int pp_tensor_potential_multiply(int rank, const int* storage_ids,
                                 const pp_cplx* potential, const int* potential_extents,
                                 double scaling,
                                 const pp_cplx* source, const int* source_extents,
                                 pp_cplx* dest, const int* dest_extents,
                                 const int32_t* const* pairs, const int* pair_counts,
                                 pp_stream stream);

/* Prepared multiply plan (replaces TensorPotential.SetupStep's lookup of
 * core.TensorPotentialMultiply_<ids> plus the PotentialData it multiplies,
 * pyprop/tensorpotential/TensorPotential.py:80-112).  The potential is copied into a
 * device layout chosen by the plan (diagonal-major bands, real storage when every
 * imaginary part is zero); the caller's `potential` buffer may be freed afterwards. */
typedef struct pp_tensor_potential pp_tensor_potential;

int pp_tp_create(int rank, const int* storage_ids, const pp_cplx* potential,
                 const int* potential_extents, const int* source_extents,
                 const int32_t* const* pairs, const int* pair_counts,
                 pp_stream stream, pp_tensor_potential** out);
/* dest += scaling * V * source (TensorPotential.MultiplyPotential, TensorPotential.py:92-112) */
int pp_tp_multiply(const pp_tensor_potential* tp, double scaling, const pp_cplx* source,
                   pp_cplx* dest, pp_stream stream);
/* Single-precision variant on complex64 vectors: dest += scaling * V * source, with V
 * held in a float32 copy made on first use and FP32 accumulation (no reference
 * counterpart: inner solves of the mixed-precision Cayley step). */
int pp_tp_multiply_f32(const pp_tensor_potential* tp, float scaling, const pp_cplxf* source,
                       pp_cplxf* dest, pp_stream stream);
/* Bytes of potential data the plan streams per multiply (for roofline accounting). */
int pp_tp_info(const pp_tensor_potential* tp, int64_t* stream_bytes, int64_t* tuple_count,
               int* is_real, int* inner_kind);
int pp_tp_destroy(pp_tensor_potential* tp);

/* Set-up helper with no reference counterpart (the reference evaluates every potential
 * on the quadrature grid and projects it, TensorPotentialGenerator.py:80-274): expand a
 * separable potential in place on the device,
 *   out[a * r_count + r] = sum_t coeffs[t * p0 + a] * radial[t * r_count + r]   (real -> complex128)
 * coeffs, radial, out: device. */
int pp_expand_potential(int terms, int64_t p0, int64_t r_count, const double* coeffs, const double* radial,
                        pp_cplx* out, pp_stream stream);

/* ------------------------------------------------------------------------------
 * Potential generation on the device (TensorPotentialGenerator.GeneratePotential,
 * TensorPotentialGenerator.py:80-274: evaluate on the quadrature grid, then represent
 * each B-spline rank in the basis, last rank first, :165-230)
 * ------------------------------------------------------------------------------ */
typedef struct pp_bspline_projection pp_bspline_projection;

/* The projection of one B-spline rank onto a pair list: the state of
 * RepresentPotentialInBasisBSpline (core/tensorpotential/basis_bspline.h:10-47) for one
 * (basis, pairs, storage, differentiation).  Host inputs: quadrature weights (n_points),
 * spline values and the differentiation-th derivative tables (n_splines, n_points), zero
 * outside each spline's support [support_first, support_first + support_count), and
 * pairs (pair_count, 2) as (row, col).  Entry (row, col) is
 * BSplineGlobalOverlapIntegral(f, differentiation, row, col)
 * (core/transform/bspline/bspline.cpp:372-466): the sum over the common support of
 * w f B_i B_j^(d) with (i, j) = (min, max)(row, col), times (-1)^d when row > col, zero
 * when |row - col| >= order, halved on the diagonal when hermitian ("Herm" storage).
 * n_points <= 2000. */
int pp_bspline_projection_create(int n_splines, int n_points, int order, const double* weights,
                                 const double* values, const double* deriv_values, const int* support_first,
                                 const int* support_count, int differentiation, int pair_count,
                                 const int* pairs, int hermitian, pp_bspline_projection** out);
int pp_bspline_projection_destroy(pp_bspline_projection* p);
int pp_bspline_projection_info(const pp_bspline_projection* p, int* n_points, int* pair_count,
                               int* max_support);
/* One rank of RepresentPotentialInBasisBSpline on the device: src (real float64) viewed
 * as (pre, n_points, post), dst as (pre, pair_count, post), float64 (dst_complex = 0) or
 * complex128 with zero imaginary part (dst_complex = 1); accumulate: dst += instead of
 * dst =.  src and dst are device buffers and must not overlap. */
int pp_bspline_represent(const pp_bspline_projection* proj, int64_t pre, int64_t post, const double* src,
                         void* dst, int dst_complex, int accumulate, pp_stream stream);
/* Grid functions (device buffers), grid[a][q1][q2] for a < pair_count:
 *   multipole: sum_l coeffs[l][a] r<^lams[l] / r>^(lams[l]+1), r< = min(r1[q1], r2[q2]),
 *     lams ascending: the 1/r12 expansion CustomPotentialEvaluationR12 evaluates
 *     (examples/tensor/helium_stabilization/coupledspherical.cpp:60-111);
 *   separable: sum_t coeffs[t][a] f1[t][q1] f2[t][q2] (one-electron potentials,
 *     examples/tensor/helium_stabilization/potential.cpp:9-76, coupledlength.cpp:9-123). */
int pp_potential_grid_multipole(int64_t pair_count, int nlam, const int* lams, const double* coeffs, int n1,
                                const double* r1, int n2, const double* r2, double* grid, pp_stream stream);
int pp_potential_grid_separable(int64_t pair_count, int terms, const double* coeffs, int n1, const double* f1,
                                int n2, const double* f2, double* grid, pp_stream stream);
/* Grid evaluation and both projections in one pass (the grid is never stored): dst
 * (pair_count, P1, P2) complex128 (+)= the potential whose grid function is the
 * multipole / separable sum above, proj1 for the first radial rank (r1, n1 = its points),
 * proj2 for the last (r2).  work: caller-owned device buffer of pair_count * n1 * P2
 * doubles (the last rank projected, first rank still on the grid).  Separable with
 * proj2 = NULL: one radial rank, dst (pair_count, P1) (+)= projection of
 * sum_t coeffs[t][a] f1[t][q] (f2 and work unused). */
int pp_potential_represent_multipole(const pp_bspline_projection* proj1, const pp_bspline_projection* proj2,
                                     int64_t pair_count, int nlam, const int* lams, const double* coeffs,
                                     const double* r1, const double* r2, double* work, pp_cplx* dst,
                                     int accumulate, pp_stream stream);
int pp_potential_represent_separable(const pp_bspline_projection* proj1, const pp_bspline_projection* proj2,
                                     int64_t pair_count, int terms, const double* coeffs, const double* f1,
                                     const double* f2, double* work, pp_cplx* dst, int accumulate,
                                     pp_stream stream);

/* ------------------------------------------------------------------------------
 * Overlap (banded Cholesky of the B-spline overlap)
 * ------------------------------------------------------------------------------ */

/* Factor the banded overlap of one B-spline rank: s_upper is host (n, sd+1) with
 * element (i, j), i <= j <= i+sd, at [j][sd + i - j] (BLAS upper Hermitian band as
 * OverlapMatrix::Setup fills it, core/representation/overlapmatrix.cpp:9-37; the
 * factorization is zpbtrf('U')).  Real-valued (the B-spline overlap is real). */
typedef struct pp_overlap pp_overlap;
int pp_overlap_create(int n, int superdiagonals, const double* s_upper, pp_overlap** out);
int pp_overlap_get_factor(const pp_overlap* ov, double* u_upper /* host (n, sd+1) */);
int pp_overlap_destroy(pp_overlap* ov);

/* Apply to every fibre of `rank` of a C-ordered tensor with `ndim` extents, in place
 * (OverlapMatrix::*Tensor loops, overlapmatrix.cpp:155-229):
 *   op 0 multiply   x <- U^H U x            (MultiplyOverlapVector, :123-128)
 *   op 1 solve      x <- U^-1 U^-H x        (SolveOverlapVector, :130-135)
 *   op 2 sqrt mul   x <- U x  | U^H x       (MultiplySqrtOverlapVector, :137-142)
 *   op 3 sqrt solve x <- U^-1 x | U^-H x    (SolveSqrtOverlapVector, :144-149)
 * conj selects the ^H form for ops 2 and 3. */
enum { PP_OVERLAP_MULTIPLY = 0, PP_OVERLAP_SOLVE = 1, PP_OVERLAP_SQRT_MULTIPLY = 2,
       PP_OVERLAP_SQRT_SOLVE = 3 };
int pp_overlap_apply(const pp_overlap* ov, int op, int conj, int ndim, const int* extents,
                     int rank, pp_cplx* data, pp_stream stream);

/* ------------------------------------------------------------------------------
 * Combined representation: the wavefunction's per-rank overlaps
 * (core/representation/combinedrepresentation.cpp:165-238, 288-481)
 * ------------------------------------------------------------------------------ */
typedef struct pp_repr pp_repr;
/* overlaps[r] = NULL for an orthogonal rank (IsOrthogonalBasis, representation.h:189-192) */
int pp_repr_create(int ndim, const int* extents, pp_overlap* const* overlaps, pp_repr** out);
int pp_repr_destroy(pp_repr* repr);
/* MultiplyOverlap / SolveOverlap / MultiplySqrtOverlap / SolveSqrtOverlap over all
 * non-orthogonal ranks in rank order, in place */
int pp_repr_overlap(const pp_repr* repr, int op, int conj, pp_cplx* psi, pp_stream stream);
/* <a|b> = a^H (S x ... x S) b (InnerProduct Algorithm 4, combinedrepresentation.cpp:423-474);
 * `work` is a caller-provided device vector of the wavefunction size; result on host. */
int pp_repr_inner_product(const pp_repr* repr, const pp_cplx* a, const pp_cplx* b,
                          pp_cplx* work, pp_cplx* result, pp_stream stream);
/* The same inner product with no device allocation (no counterpart in the reference): for
 * callers that cannot allocate (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Every intermediate lives in a fixed file-scope device array of
 * 2 PP_NOALLOC_MAX_ELEMENTS + 1025 complex128 values (32 MiB) per device, held until the
 * process ends; wavefunctions larger than PP_NOALLOC_MAX_ELEMENTS are rejected.  Calls that
 * overlap in time on one device share that array and race: serialise them.  Same kernels
 * and result as pp_repr_inner_product, bitwise. */
#define PP_NOALLOC_MAX_ELEMENTS (1LL << 20)
int pp_repr_inner_product_noalloc(const pp_repr* repr, const pp_cplx* a, const pp_cplx* b, pp_cplx* result,
                                  pp_stream stream);

/* ------------------------------------------------------------------------------
 * Hamiltonian: sum of tensor potentials (BasisPropagator,
 * pyprop/tensorpotential/BasisPropagator.py:116-159)
 * ------------------------------------------------------------------------------ */
typedef struct pp_hamiltonian pp_hamiltonian;
/* Fuses the given plans into one launch where their storage allows.  The plans stay
 * owned by the caller and must outlive the Hamiltonian. */
int pp_ham_create(const pp_repr* repr, int count, pp_tensor_potential* const* potentials,
                  pp_hamiltonian** out);
int pp_ham_destroy(pp_hamiltonian* ham);
/* dest = sum_p scalings[p] V_p src      (MultiplyHamiltonianNoOverlap, :126-129) */
int pp_ham_multiply_nooverlap(pp_hamiltonian* ham, const double* scalings,
                              const pp_cplx* src, pp_cplx* dest, pp_stream stream);
/* dest = S^-1 sum_p scalings[p] V_p src (MultiplyHamiltonian, :116-124) */
int pp_ham_multiply(pp_hamiltonian* ham, const double* scalings, const pp_cplx* src,
                    pp_cplx* dest, pp_stream stream);
/* dest = U^-H (sum_p scalings[p] V_p) U^-1 src, src unchanged
 * (MultiplyHamiltonianBalancedOverlap, :131-159) */
int pp_ham_multiply_balanced(pp_hamiltonian* ham, const double* scalings,
                             const pp_cplx* src, pp_cplx* dest, pp_stream stream);

/* ------------------------------------------------------------------------------
 * BLAS-1 on device vectors (piram functors / blitzblas,
 * core/krylov/piram/piram/functors.h:185-241; blitzblas.h:379-487)
 * Results returned to host; deterministic reduction order.
 * ------------------------------------------------------------------------------ */
int pp_vec_dot(int64_t n, const pp_cplx* x, const pp_cplx* y, pp_cplx* result, pp_stream stream); /* sum conj(x) y */
int pp_vec_norm(int64_t n, const pp_cplx* x, double* result, pp_stream stream);                  /* sqrt(Re <x,x>) */

/* ------------------------------------------------------------------------------
 * Krylov solvers.  An operator is a function table: out = A in on `stream`.
 * ------------------------------------------------------------------------------ */
typedef int (*pp_operator_fn)(void* ctx, const pp_cplx* in, pp_cplx* out, pp_stream stream);

/* GMRES (core/krylov/gmres/gmres.h:28-525; wrapper core/krylov/gmres/gmreswrapper.cpp):
 * single-cycle, x0 = 0, absolute residual tolerance clamped to >= 1e-13, optional
 * DGKS re-orthogonalisation (threshold 0.717, at most 5 passes). */
typedef struct pp_gmres pp_gmres;
int pp_gmres_create(int64_t n, int basis_size, double tolerance, int double_orthogonalization,
                    pp_gmres** out);
int pp_gmres_destroy(pp_gmres* g);
/* Solve A x = rhs; rhs is preserved; x overwritten (GMRES::SolveVector, gmres.h:433-505). */
int pp_gmres_solve(pp_gmres* g, pp_operator_fn op, void* ctx, const pp_cplx* rhs, pp_cplx* x,
                   pp_stream stream);
/* Error estimate of the last solve, operator applications, Arnoldi steps
 * (GetErrorEstimate, GetOperatorCount; gmres.h:161-194). */
int pp_gmres_stats(const pp_gmres* g, double* error_estimate, int* operator_count,
                   int* orthogonalization_count);
/* Error estimate per Arnoldi step of the last solve (GetErrorEstimateList); `out`
 * holds basis_size doubles. */
int pp_gmres_error_list(const pp_gmres* g, double* out);

/* pAMP (core/krylov/pamp/pamp/pamp.h:28-690; wrapper pampwrapper.cpp): Arnoldi with
 * classical Gram-Schmidt (+ optional re-orthogonalisation, threshold 0.717, at most 3
 * passes), scaling-and-squaring Pade exponential of the Hessenberg matrix, all on the
 * device (no host synchronisation inside a step). */
typedef struct pp_pamp pp_pamp;
int pp_pamp_create(int64_t n, int basis_size, int pade_order, int double_orthogonalization,
                   pp_pamp** out);
int pp_pamp_destroy(pp_pamp* p);
/* exponentiation of the Hessenberg matrix (krylov_exponentiation_method, pampwrapper.cpp:56-59):
   0 = Pade (the default, on the device), 1 = eigenvector exponential (pamp.h:567-608: the
   eigen-decomposition of the m x m matrix on the host, one synchronisation per step). */
int pp_pamp_set_exponentiation(pp_pamp* p, int method);
/* v <- exp(-i dt A) v (pAMP::PropagateVector, pamp.h:613-670).  Must not be given an
 * operator that synchronises the stream if the caller wants to capture the step. */
int pp_pamp_propagate(pp_pamp* p, pp_operator_fn op, void* ctx, pp_cplx* v, pp_cplx dt,
                      pp_stream stream);
/* |exp(H)[m-1,0]| of the last step (GetPropagationErrorEstimate) and happy-breakdown
 * flag; synchronises the stream. */
int pp_pamp_stats(const pp_pamp* p, pp_cplx* error_estimate, int* breakdown, int* operator_count);

/* ------------------------------------------------------------------------------
 * Block-diagonal ("radial") preconditioner
 * (examples/tensor/helium_stabilization/preconditioner.py:77-198)
 * ------------------------------------------------------------------------------ */
typedef struct pp_precond pp_precond;
/* Exact factorization of M_c = scaling_s * (S1 x S2) + scaling_h * P_c for every angular
 * channel c of a rank-3 wavefunction (nchan, n1, n2).
 *   s1, s2: host dense (n1, n1), (n2, n2) overlaps; band = half bandwidth of S (k-1).
 *   p_data: host complex (nchan, p1_count, p2_count) angular-diagonal potential with
 *   Simp pair lists pairs1 (p1_count, 2) and pairs2 (p2_count, 2) for the radial ranks.
 * LU without pivoting (the real part of M_c is positive definite); the SuperLU variant
 * of the reference solves the same systems exactly. */
int pp_precond_create(int nchan, int n1, int n2, int band, const double* s1, const double* s2,
                      const pp_cplx* p_data, int p1_count, const int32_t* pairs1, int p2_count,
                      const int32_t* pairs2, pp_cplx scaling_h, pp_cplx scaling_s,
                      pp_precond** out);
int pp_precond_destroy(pp_precond* pc);
/* psi <- M^-1 psi per channel (Preconditioner.Solve, preconditioner.py:189-198) */
int pp_precond_solve(const pp_precond* pc, pp_cplx* psi, pp_stream stream);

/* ------------------------------------------------------------------------------
 * Cayley / Crank-Nicolson step (pyprop/propagator/CayleyPropagator.py:71-150)
 *   (S + i dt/2 H(t+dt/2)) psi(t+dt) = (S - i dt/2 H(t+dt/2)) psi(t)
 * left-preconditioned by M (optional), solved by pp_gmres.
 * ------------------------------------------------------------------------------ */
typedef struct pp_cayley pp_cayley;
int pp_cayley_create(pp_hamiltonian* ham, const pp_precond* pc, int basis_size, double tolerance,
                     int double_orthogonalization, pp_cayley** out);
int pp_cayley_destroy(pp_cayley* c);
/* scalings: per-potential time factors at t + |dt|/2 */
int pp_cayley_step(pp_cayley* c, pp_cplx* psi, double dt, const double* scalings,
                   pp_stream stream);
int pp_cayley_stats(const pp_cayley* c, double* error_estimate, int* operator_count);

/* pAMP step in the balanced-overlap form (PamPropagator.AdvanceStep,
 * pyprop/propagator/PamPropagator.py:50-59): psi <- U^-1 exp(-i dt U^-H H U^-1) U psi */
int pp_pamp_step_balanced(pp_pamp* p, pp_hamiltonian* ham, pp_cplx* psi, pp_cplx dt,
                          const double* scalings, pp_stream stream);

/* ------------------------------------------------------------------------------
 * Kronecker (separable) preconditioner.  No reference counterpart: a replacement for
 * the block-diagonal RadialTwoElectronPreconditioner
 * (examples/tensor/helium_stabilization/preconditioner.py:101-198) that inverts the
 * separable part of each channel block exactly,
 *   K_c = S (x) S + a (h_l1 (x) S + S (x) h_l2),
 * from the generalized eigenpairs h_l E_l = S E_l diag(lam_l), E_l^T S E_l = I:
 *   K_c^-1 B = E_l1 [(E_l1^T B E_l2) / (1 + a (lam_l1,i + lam_l2,j))] E_l2^T.
 * S: host [n][n] radial overlap, E: host [nl][n][n] row-major (column j = eigenvector
 * j), lam: host [nl][n], l1, l2: host [nchan] (the l index of each electron in channel
 * c).  psi: [nchan][n][n].
 * ------------------------------------------------------------------------------ */
typedef struct pp_kron pp_kron;
int pp_kron_create(int nchan, int n, int nl, const double* S, const double* E, const double* lam,
                   const int* l1, const int* l2, pp_kron** out);
int pp_kron_destroy(pp_kron* k);
/* psi <- K^-1 psi (complex128 / complex64) */
int pp_kron_solve(const pp_kron* k, pp_cplx a, pp_cplx* psi, pp_stream stream);
/* psi <- K psi (complex128) */
int pp_kron_multiply(const pp_kron* k, pp_cplx a, pp_cplx* psi, pp_stream stream);
int pp_kron_solve_f32(const pp_kron* k, pp_cplx a, pp_cplxf* psi, pp_stream stream);
/* out <- K^-1 in in FP32 (no reference counterpart; in and out distinct, so the one-launch
 * form can run). */
int pp_kron_solve_f32_into(const pp_kron* k, pp_cplx a, const pp_cplxf* in, pp_cplxf* out, pp_stream stream);
/* The same, with sz <- (S (x) S) out formed from the solve's middle product G as F1 G F2^T
 * (no reference counterpart; the one-launch form only, which this requires). */
int pp_kron_solve_f32_overlap(const pp_kron* k, pp_cplx a, const pp_cplxf* in, pp_cplxf* out, pp_cplxf* sz,
                              pp_stream stream);
/* The FP32 solve's form for n <= 64 (no reference counterpart): 0 one workgroup per channel,
 * 1 four workgroups per channel in two launches, 2 (the default) four per channel in one launch
 * (each forms the whole middle product; used when input, output and write-back are distinct,
 * else form 1).  Forms 1 and 2 give the same bits. */
int pp_kron_set_form(pp_kron* k, int form);
/* use K (a = i dt/2) as the left preconditioner of a Cayley solver created without a
 * block preconditioner (NULL detaches) */
int pp_cayley_set_kron(pp_cayley* c, const pp_kron* k);

/* ------------------------------------------------------------------------------
 * Learned preconditioner.  No reference counterpart.  An MLP shared by all channel
 * blocks (in = 2 n^2 reals, hidden width `hidden`, GELU, per-channel embedding added to
 * the first hidden layer), BF16 MFMA products with FP32 accumulation, FP32 master
 * weights trained by Adam on the device.  As a preconditioner it applies
 * K_c^-1 (v_c + sigma |v| net(v_c / |v|)), |v| the norm of the whole vector.
 * Parameter vector layout (FP32): W1 [hidden][in], b1 [hidden], emb [nchan][hidden],
 * W2 [hidden][hidden], b2 [hidden], W3 [in][hidden], b3 [in].
 * x, t, y: device FP32 [rows][in] (block c as n*n interleaved re, im); chan: device
 * int32 [rows].
 * ------------------------------------------------------------------------------ */
typedef struct pp_learned pp_learned;
int pp_learned_create(int nchan, int n, int hidden, unsigned long long seed, pp_learned** out);
int pp_learned_destroy(pp_learned* l);
int pp_learned_param_count(const pp_learned* l, long long* count);
int pp_learned_get_params(const pp_learned* l, float* host);
int pp_learned_set_params(pp_learned* l, const float* host);
int pp_learned_get_grads(const pp_learned* l, float* host);
int pp_learned_set_scale(pp_learned* l, float sigma);
/* y = net(x) (before sigma) */
int pp_learned_forward(pp_learned* l, int rows, const float* x, const int* chan, float* y,
                       pp_stream stream);
/* gradient of loss = sum |net(x) - t / sigma|^2 / rows into the parameter gradient;
 * loss (host, may be NULL) */
int pp_learned_gradient(pp_learned* l, int rows, const float* x, const int* chan, const float* t,
                        double* loss, pp_stream stream);
/* gradient + one Adam step (beta 0.9, 0.999, eps 1e-8) */
int pp_learned_train_step(pp_learned* l, int rows, const float* x, const int* chan, const float* t,
                          float lr, double* loss, pp_stream stream);
/* training rows from one solved system A d = r (complex128 psi-shaped r, d): per channel c
 * x_c = r_c / |r|, t_c = (K_c d_c - r_c) / |r|, nrm_c = |r_c| / |r| (device FP32 outputs,
 * x and t [nchan][in], nrm [nchan]) */
int pp_learned_prepare(pp_learned* l, const pp_kron* k, double dt, const pp_cplx* r, const pp_cplx* d,
                       float* x_out, float* t_out, float* nrm_out, pp_stream stream);
/* out = K^-1 (in + sigma |in| net(in_c / |in|)) (complex64, psi-shaped) */
int pp_learned_apply(pp_learned* l, const pp_kron* k, double dt, const pp_cplxf* in, pp_cplxf* out,
                     pp_stream stream);
/* C[M][N] = sum_k A(m,k) B(k,n), A(m,k) = a[m*sam + k*sak], B(k,n) = b[k*sbk + n*sbn]
 * (BF16 bit patterns, FP32 result, MFMA) */
int pp_gemm_bf16(int M, int N, int K, const uint16_t* a, long long sam, long long sak,
                 const uint16_t* b, long long sbk, long long sbn, float* c, pp_stream stream);

/* ------------------------------------------------------------------------------
 * Factored Hamiltonian multiply.  No reference counterpart for the evaluation order: the
 * operator is the sum of the tensor potentials (BasisPropagator.MultiplyHamiltonianNoOverlap,
 * BasisPropagator.py:116-159) written as sum_t f_p(t) (A_t (x) R_t), one angular matrix A_t
 * (nchan x nchan, host [T][nchan][nchan]) and one radial operator R_t per separable term
 * (the multipole sum of 1/r12, coupledspherical.cpp:60-111, is 11 such terms), R_t given
 * diagonal-blocked as host dia[t][i1][jj][d][i2] = element (row (i1, i2), column
 * (i1-bw+jj, i2-bw+d)).  term_potential[t] indexes the per-potential scalings.  The angular
 * contraction W_t = f A_t X and the radial products both run on MFMA (FP64 and FP32).
 * psi: [nchan][n][n].
 * ------------------------------------------------------------------------------ */
typedef struct pp_fham pp_fham;
int pp_fham_create(int nchan, int n, int nterms, int bw, const double* A, const int* term_potential,
                   const double* dia, int npotentials, pp_fham** out);
int pp_fham_destroy(pp_fham* f);
/* Terms whose band is an outer product a (x) b run as two 1-D band passes instead of the
 * MFMA band tiles; band_terms receives the count left on the tiles (nterms when none is
 * separable). */
int pp_fham_band_terms(const pp_fham* f, int* band_terms);
/* dst = addend + alpha * sum_p scalings[p] V_p src (addend may be NULL; src may alias dst) */
int pp_fham_multiply(const pp_fham* f, const double* scalings, const pp_cplx* src, pp_cplx* dst,
                     const pp_cplx* addend, pp_cplx alpha, pp_stream stream);
int pp_fham_multiply_f32(const pp_fham* f, const double* scalings, const pp_cplxf* src, pp_cplxf* dst,
                         const pp_cplxf* addend, pp_cplx alpha, pp_stream stream);
/* Radial-stage kernel shape of the factored multiply (no reference counterpart: the
 * reference has no factored form).  Variants 0..6 multiply band tiles: (terms x groups of 8
 * channels) = (1,1), (2,2), (3,2), (3,3) per wave, and 4..6 the shapes of 1..3 with the next
 * band step's operands prefetched into registers.  Variants 7..9 form each position's
 * products of all band terms as one small GEMM over the w x w window (no band padding), with
 * 6, 3 or 2 channel groups per wave; they need the separable split (n <= 64) and
 * pp_fham_tune reports -1 for them without it.  precision is 64 (complex128 multiply) or 32
 * (complex64).  pp_fham_tune times every radial variant (trials multiplies each, on stream)
 * and keeps the fastest, then does the same for the angular stage's form: variants 0..2 run
 * one row block of 16 channels per workgroup with 16, 8 or 4 waves splitting its K, 3..4 all
 * row blocks of a 16-position tile per workgroup (each Z element read once; up to 64
 * channels) with 8 or 16 waves.  timings_ms (may be NULL) receives PP_FHAM_RADIAL_VARIANTS
 * radial then PP_FHAM_ANGULAR_VARIANTS angular entries.  set_angular_waves selects among
 * variants 0..2 by wave count; get_angular_waves returns the wave count of the current
 * variant. */
#define PP_FHAM_RADIAL_VARIANTS 10
#define PP_FHAM_ANGULAR_VARIANTS 5
int pp_fham_set_radial_variant(pp_fham* f, int precision, int variant);
int pp_fham_get_radial_variant(const pp_fham* f, int precision, int* variant);
int pp_fham_set_angular_waves(pp_fham* f, int precision, int waves);
int pp_fham_get_angular_waves(const pp_fham* f, int precision, int* waves);
int pp_fham_set_angular_variant(pp_fham* f, int precision, int variant);
int pp_fham_get_angular_variant(const pp_fham* f, int precision, int* variant);
int pp_fham_tune(pp_fham* f, int precision, int trials, int* best, double* timings_ms, pp_stream stream);
/* Overlap term (no reference counterpart): appends S1 (x) S2 (dense n x n, row-major, zero
 * outside the operator's band) as one more radial term whose angular part is the identity,
 * so that one multiply forms the Cayley operator
 *   dst = addend + alpha * sum_p scalings[p] V_p src + beta * (S1 (x) S2) src
 * (CayleyPropagator.py:116-142 builds the same S + i dt/2 H from MultiplyOverlap and the
 * potentials).  The overlap radial product is only formed when beta != 0. */
int pp_fham_set_overlap(pp_fham* f, const double* S1, const double* S2);
int pp_fham_multiply_overlap(const pp_fham* f, const double* scalings, const pp_cplx* src, pp_cplx* dst,
                             const pp_cplx* addend, pp_cplx alpha, pp_cplx beta, pp_stream stream);
int pp_fham_multiply_overlap_f32(const pp_fham* f, const double* scalings, const pp_cplxf* src, pp_cplxf* dst,
                                 const pp_cplxf* addend, pp_cplx alpha, pp_cplx beta, pp_stream stream);
/* use the factored multiply inside a Cayley solver (NULL detaches) */
int pp_cayley_set_factored(pp_cayley* c, const pp_fham* f);

/* ------------------------------------------------------------------------------
 * Initial-state eigensolvers.
 *
 * pIRAM: implicitly restarted Arnoldi (core/krylov/piram/piram/piram.h:28-690, wrapped by
 * pyprop/solver/PiramSolver.py:3-141).  basis_size = krylov_basis_size (>= 2 x
 * eigenvalue_count, piram.h:269-272), eigenvalue_count = krylov_eigenvalue_count,
 * max_restarts = krylov_max_iteration_count.  The wanted Ritz values are those with the
 * smallest real part (CompareComplexLessReal, functors.h:95-102), shifts are the others.
 * Convergence: |y_m| |f| <= tolerance max(eps^(2/3), |theta|) (ARPACK's test; see
 * eigen.hip for the reference's variant).  start = NULL gives the reference's uniform
 * [0,1) real start (fixed seed).  op is any pp_operator_fn: pp_ham_op_apply (the
 * Hamiltonian), pp_shift_invert_apply (inverse iterations) or a caller's own.
 * ------------------------------------------------------------------------------ */
typedef struct pp_piram pp_piram;
int pp_piram_create(int64_t n, int basis_size, int eigenvalue_count, double tolerance, int max_restarts,
                    int max_orthogonalization, pp_piram** out);
int pp_piram_destroy(pp_piram* p);
int pp_piram_solve(pp_piram* p, pp_operator_fn op, void* ctx, const pp_cplx* start, pp_stream stream);
/* GetEigenvalues / GetErrorEstimates / GetConvergenceEstimates (PiramSolver.py:116-121,
 * piram.h:182-199): eigenvalue_count values, ascending real part */
int pp_piram_eigenvalues(const pp_piram* p, pp_cplx* values, double* error_estimates,
                         double* convergence_estimates, int* converged_count);
/* GetEigenvector (PiramSolver.py:123-134): unit 2-norm Ritz vector into out (device) */
int pp_piram_eigenvector(pp_piram* p, int index, pp_cplx* out, pp_stream stream);
int pp_piram_stats(const pp_piram* p, int* operator_count, int* restart_count, int* orthogonalization_count);

/* GMRESShiftInvertSolver (pyprop/solver/GMRESShiftInvertSolver.py:1-132): dst = (H - shift S)^-1 S src
 * by GMRES(basis_size, tolerance) with left preconditioner pc (may be NULL), which the caller
 * sets up with Hamiltonian scaling 1 and overlap scaling -shift (as __init__ does, 44-46).
 * pp_shift_invert_apply has the pp_operator_fn signature (ctx = the pp_shift_invert*), the
 * InverseIterations matrix_vector_func that PiramSolver takes (PiramSolver.py:62-70). */
typedef struct pp_shift_invert pp_shift_invert;
int pp_shift_invert_create(pp_hamiltonian* ham, const pp_precond* pc, int basis_size, double tolerance,
                           int double_orthogonalization, double shift, pp_shift_invert** out);
int pp_shift_invert_destroy(pp_shift_invert* si);
int pp_shift_invert_set_scalings(pp_shift_invert* si, const double* scalings);
int pp_shift_invert_set_factored(pp_shift_invert* si, const pp_fham* f);
int pp_shift_invert_solve(pp_shift_invert* si, const pp_cplx* src, pp_cplx* dst, pp_stream stream);
int pp_shift_invert_apply(void* ctx, const pp_cplx* in, pp_cplx* out, pp_stream stream);
/* negate != 0: pp_shift_invert_apply returns -(H - shift S)^-1 S.  With a shift below the
 * spectrum the Ritz values -1 / (E - shift) are then all negative and pIRAM's wanted set
 * (smallest real part) holds the lowest energies, lowest first (E = shift - 1 / theta).
 * pp_shift_invert_solve is not affected. */
int pp_shift_invert_set_negate(pp_shift_invert* si, int negate);
int pp_shift_invert_stats(const pp_shift_invert* si, int* solves, int* operator_count, double* last_error);

/* The Hamiltonian as a pp_operator_fn (ctx = the pp_ham_op*): mode 0 = H
 * (MultiplyHamiltonianNoOverlap), 1 = S^-1 H (MultiplyHamiltonian), 2 = U^-T H U^-1
 * (MultiplyHamiltonianBalancedOverlap), BasisPropagator.py:116-159; scalings fixed at
 * creation (NULL: all 1). */
typedef struct pp_ham_op pp_ham_op;
int pp_ham_op_create(pp_hamiltonian* ham, const double* scalings, int mode, pp_ham_op** out);
int pp_ham_op_destroy(pp_ham_op* h);
int pp_ham_op_apply(void* ctx, const pp_cplx* in, pp_cplx* out, pp_stream stream);

/* ------------------------------------------------------------------------------
 * Mixed-precision Cayley step.  No reference counterpart; solves the system of
 * pp_cayley_step (CayleyPropagator.py:71-150) by FP64 iterative refinement around a
 * complex64 flexible GMRES with the Kronecker preconditioner (plus the learned
 * correction on the first direction when attached):
 *   stop when |b - A x| <= tol_rel |b| (FP64 residual); inner solves (basis_size
 *   1..60) stop at inner_rtol relative to their right-hand side.  Each correction runs
 *   as many inner iterations as it needed in the previous step before the host reads
 *   any result (more columns than needed only lower the residual further).
 * With a fallback solver attached, a step whose inner iterations exceed
 * fallback_iterations (or that does not converge in max_outer corrections) is redone
 * by the plain solver from the unchanged psi.
 * counts (7 ints): outer corrections, inner iterations, FP64 multiplies, FP32
 * multiplies, preconditioner applications, learned applications, fallback flag;
 * residuals (2 doubles): |b - A x|, |b - A x| / |b|.
 * ------------------------------------------------------------------------------ */
typedef struct pp_mixed pp_mixed;
int pp_mixed_create(pp_hamiltonian* ham, const pp_kron* k, int basis_size, double tol_rel,
                    double inner_rtol, int max_outer, pp_mixed** out);
int pp_mixed_destroy(pp_mixed* m);
int pp_mixed_set_learned(pp_mixed* m, pp_learned* l);
/* where the learned correction acts (flexible GMRES allows a different preconditioner per
 * direction): 0 = the first direction of the first correction (default), 1 = the first
 * direction of every correction, 2 = every direction */
int pp_mixed_set_learned_mode(pp_mixed* m, int mode);
int pp_mixed_set_fallback(pp_mixed* m, pp_cayley* plain, int fallback_iterations);
/* use the factored multiply for the FP64 residuals and the FP32 inner products (NULL detaches) */
int pp_mixed_set_factored(pp_mixed* m, const pp_fham* f);
/* device_mode 1: once a step has converged, the next steps run as many corrections and
 * inner iterations as it needed with every inner least-squares problem solved on the
 * device, and synchronise with the host once, at the final residual (a step not converged
 * then continues host-driven).  0 (the default): every correction is host-driven, with
 * the inner iteration count adapted at its synchronisations.  The results agree; on one
 * MI355X with helium both modes take 1.41 ms per step (the step is kernel-bound). */
int pp_mixed_set_device_mode(pp_mixed* m, int device_mode);
/* Classical Gram-Schmidt passes per inner FGMRES iteration (no reference counterpart: the
 * reference's GMRES orthogonalises in FP64, gmres.h:433-505).  1 (the default): one pass,
 * the norm taken from the updated vector (3 launches); the FP32 basis is then orthogonal
 * only to ~eps32 / h_{j+1,j}, which the FP64 refinement absorbs (at helium: the same 3 inner
 * iterations per correction and 1e-11 error after 20 steps).  2: two passes (4 launches). */
int pp_mixed_set_gram_schmidt(pp_mixed* m, int passes);
/* Where a one-pass Gram-Schmidt is finished (no reference counterpart): 1 (the default) by the
 * next Kronecker solve when the host reads nothing before it (the solve sums the block
 * partials itself; one launch fewer per such iteration, bitwise the same results); 0 by a
 * finishing launch of its own after every iteration. */
int pp_mixed_set_deferred_finish(pp_mixed* m, int on);
/* The FP32 overlap product of each preconditioner output z (no reference counterpart): 1 (the
 * default) formed by the Kronecker solve that produced z, as F1 G F2^T (F_l = S E_l, G the
 * solve's middle product; E^T S E = I makes it S1 z S2), so the FP32 multiply drops its
 * Cholesky-form overlap passes; 0 by those passes, U1^T U1 z U2^T U2.  Only with the one-launch
 * Kronecker form and the factored operator's Cholesky factors attached. */
int pp_mixed_set_overlap_from_solve(pp_mixed* m, int on);
/* Graph mode (no reference counterpart; with device mode 1): a device-mode step is one fixed
 * sequence of launches for given correction and inner-iteration counts; it is captured into
 * a hipGraph once, after the same sequence has run eagerly, and replayed on the caller's
 * stream for every later step with the same counts (recaptured when they change).  The time
 * scalings are staged in device memory each step.  Needs the factored operator with a sparse
 * angular form in both precisions (the tuned form at helium); otherwise the step runs
 * eagerly.  Results are bitwise those of eager device mode. */
int pp_mixed_set_graph(pp_mixed* m, int graph_mode);
int pp_mixed_graph_stats(const pp_mixed* m, int* captures, int* launches);
int pp_mixed_step(pp_mixed* m, pp_cplx* psi, double dt, const double* scalings, pp_stream stream);
int pp_mixed_stats(const pp_mixed* m, int* counts, double* residuals);

/* ------------------------------------------------------------------------------
 * Batched pAMP propagation of independent atoms.  No reference counterpart for the
 * batching; each atom's step is PamPropagator.AdvanceStep in the balanced-overlap form
 * (pyprop/propagator/PamPropagator.py:50-59, pAMP::PropagateVector pamp.h:270-670) with the
 * reduced-spherical hydrogen Hamiltonian H0 (Diag x BandNH) + E(t) dipole (Simp x BandNH,
 * couplings l -> l +- 1).  One persistent launch advances every atom by one step (one
 * workgroup per atom, largest first); steps can be captured into a hipGraph and replayed.
 * Per-atom host arrays are concatenated in atom order:
 *   h0:     [nl][2b+1][N]   H0 in DIA form, element (i, i+d-b) at [l][d][i]
 *   r:      [2b+1][N]       the dipole's radial matrix in DIA form
 *   cpl:    [nl][2]         angular dipole couplings to l-1 and l+1
 *   s_upper:[N][b+1]        radial overlap, upper band (element (i, j) at [j][b+i-j])
 *   pulse:  [3]             E0, omega, duration of E0 sin^2(pi t/T) cos(omega t)
 * fp32_matvec: the banded H multiply in complex64 with FP32 accumulation (state, Krylov
 * basis, orthogonalisation, Pade and overlap solves stay FP64).
 * ------------------------------------------------------------------------------ */
typedef struct pp_batch pp_batch;
int pp_batch_create(int natoms, const int* nl, const int* n, int band, const double* h0, const double* r,
                    const double* cpl, const double* s_upper, const double* pulse, int krylov, int pade_order,
                    int fp32_matvec, pp_batch** out);
int pp_batch_destroy(pp_batch* b);
/* host psi (concatenated [nl][N] complex128 per atom) and per-atom times; waits for the
 * device first (a step or replay on any stream finishes before the state is replaced) */
int pp_batch_set_state(pp_batch* b, const pp_cplx* psi, const double* t);
/* any of psi, t, err (per-atom error estimate E(m-1, 0)) may be NULL */
int pp_batch_get_state(const pp_batch* b, pp_cplx* psi, double* t, pp_cplx* err);
/* time 64..1024-thread workgroups on trial_steps steps, restore the state, keep the fastest;
 * timings_ms: 5 per-step times (may be NULL) */
int pp_batch_autotune(pp_batch* b, double dt, int trial_steps, int* best_threads, double* timings_ms,
                      pp_stream stream);
int pp_batch_set_threads(pp_batch* b, int threads);
/* nsteps persistent launches (one step each) */
int pp_batch_step(pp_batch* b, double dt, int nsteps, pp_stream stream);
/* capture steps_per_graph step launches into a hipGraph on a non-default stream; replay it */
int pp_batch_capture(pp_batch* b, double dt, int steps_per_graph, pp_stream stream);
int pp_batch_replay(pp_batch* b, int replays, pp_stream stream);
int pp_batch_info(const pp_batch* b, int* threads, int* grid, long long* psi_elements, long long* lds_bytes);

#ifdef __cplusplus
}
#endif

#endif /* PYPROP_AMD_H */

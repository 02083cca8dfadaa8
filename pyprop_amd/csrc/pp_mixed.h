// Mixed-precision Cayley solve: Kronecker preconditioner, FP32 flexible GMRES, FP64
// iterative refinement, and the hook for the learned preconditioner.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <functional>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"

namespace pp {

// A one-pass Gram-Schmidt finish left to the next preconditioner launch: the block partials
// of the dots (p1, rows of m + 1) and of the updated |w|^2 (p3, one per block), and where the
// finished values go (the layout of finish_cgs1_kernel)
struct GsFinish {
  const double2* p1 = nullptr;
  const double2* p3 = nullptr;
  int blocks = 0, m = 0;
  double2* out = nullptr;
  double2* out2 = nullptr;
};

// Exact inverse of the separable part of every channel block of S (x) S + a H0:
//   K_c = S x S + a (h_l1 x S + S x h_l2),   h_l E_l = S E_l diag(lam_l),  E_l^T S E_l = I.
struct Kron {
  int nchan = 0, n = 0, nl = 0;
  DeviceBuffer<double> E64, lam64;   // [l][n][n] row-major, [l][n]
  DeviceBuffer<double> F64;          // F_l = S E_l (= E_l^-T), for the forward product
  DeviceBuffer<float> E32, lam32;
  DeviceBuffer<float> F32;           // F_l in FP32: the overlap product S1 X S2 = F1 G F2^T of a solve
  DeviceBuffer<int> l1, l2;          // per channel
  mutable DeviceBuffer<float2> g32;  // [c][64][64] middle product of the split FP32 solve
  bool split32 = true;               // FP32 solve (n <= 64) with four workgroups per channel
  bool fused32 = true;               // ... in one launch (each workgroup forms the whole middle product)
  void setup_attributes();
  // psi <- K^-1 psi, a = complex factor of the one-electron part (i dt / 2)
  void apply64(double2 a, double2* psi, hipStream_t s) const;
  void apply32(double2 a, float2* psi, hipStream_t s) const;
  // out <- K^-1 (in_scale in) (may alias in); wb (may be NULL, not out): wb <- in_scale in
  // as well; in_n2 (device, may be NULL): in_scale = 1 / sqrt(in_n2->x).  fin (may be NULL):
  // a pending Gram-Schmidt finish whose |w|^2 is *in_n2: the launch performs it (the one-launch
  // form sums the partials itself; the others run the finishing kernel first)
  // sz (may be NULL): the overlap product (S (x) S) out as well, from the solve's middle
  // product (one-launch form only: see overlap_product32)
  void apply32(double2 a, const float2* in, float2* out, hipStream_t s, float in_scale, float2* wb,
               const double2* in_n2, const GsFinish* fin = nullptr, float2* sz = nullptr) const;
  // whether apply32 with these buffers runs the one-launch form, which can form sz
  bool overlap_product32(const float2* in, const float2* out, const float2* wb) const;
  // out <- K^-1 (in + add_sigma * (*add_norm) * add), *add_norm read on the device; sz (may be
  // NULL, needs overlap_product32(in, out, NULL)): also (S (x) S) out, as apply32 forms it
  void apply32_add(double2 a, const float2* in, const float2* add, const double* add_norm, float add_sigma,
                   float2* out, hipStream_t s, float2* sz = nullptr) const;
  void launch32(double2 a, const float2* in, float2* out, hipStream_t s, float in_scale, float2* wb,
                const float2* add, const double* add_norm, float add_sigma, const double2* in_n2,
                const GsFinish* fin = nullptr, float2* sz = nullptr) const;
  // psi <- K psi (FP64)
  void multiply64(double2 a, double2* psi, hipStream_t s) const;
};

struct Learned;  // learned.hip

struct MixedStats {
  int outer_iterations = 0;   // FP32 corrections applied
  int inner_iterations = 0;   // FP32 FGMRES iterations, summed
  int outer_ops = 0;          // FP64 Hamiltonian multiplies
  int inner_ops = 0;          // FP32 Hamiltonian multiplies
  int precond_applies = 0;
  int learned_applies = 0;
  int fallback = 0;           // 1 when the step was redone by the plain solver
  double residual = 0;        // final |b - A x| (FP64)
  double rel_residual = 0;    // final |b - A x| / |b|
};

struct MixedWork {
  static constexpr int kSmall = 8192;  // inner basis sizes up to 60
  long long n = -1;
  int m = 0;
  DeviceBuffer<float2> V, Z, w, tmp;
  DeviceBuffer<double2> b64, r64, x64, t64, h64;
  DeviceBuffer<float2> wnext;  // the Gram-Schmidt-updated direction of the last iteration
  DeviceBuffer<float2> sz;     // (S (x) S) z of the last preconditioner output z, when its solve formed it
  const float2* sz_for = nullptr;  // that z (NULL: sz holds nothing usable)
  double2* small = nullptr;      // kSmall results for the host (pinned, mapped, coherent)
  double2* small_dev = nullptr;  // its device address
  Reducer red;
  std::vector<double2> host_small;
  MixedWork() = default;
  MixedWork(const MixedWork&) = delete;
  MixedWork& operator=(const MixedWork&) = delete;
  ~MixedWork();
  void ensure(long long n, int m);
  void read_small(int count, hipStream_t s);  // synchronise s, copy small[0..count) to host_small
  DeviceBuffer<double2> norms;   // device copies of step-level squared norms (|b|^2, |r0|^2, |r|^2)
  DeviceBuffer<double> npart;    // per-workgroup squared norms from the FP64 angular epilogue
  DeviceBuffer<double2> hdev;    // device copy of the inner solve's results (the Arnoldi scaling)
  DeviceBuffer<float2> ycoef;    // inner least-squares solution of the device mode
  // out[0] = |a|^2, out[1] = |b|^2 (pinned), the same into out_dev (device)
  void norm2_pair(const double2* a, const double2* b, double2* out, double2* out_dev, hipStream_t s);
  // Two-pass classical Gram-Schmidt of w against V_0..V_{k-1} (FP64 sums): out[0..k] = first
  // dots and |w|^2, out[k+1..2k+1] = second dots and |w|^2, out[2k+2] = final |w|^2
  void cgs2_32(int k, float2* w, double2* out, double2* out_dev, hipStream_t s);
  // one pass: dots, then the update with the norm of the updated vector; with defer, the
  // finish is left in pend for the next preconditioner launch
  void cgs1_32(int k, float2* w, double2* out, double2* out_dev, hipStream_t s, bool defer = false);
  // defer: the caller issues the next iteration's preconditioner before anything reads the
  // results (one-pass form only)
  void gram_schmidt(int k, float2* w, double2* out, double2* out_dev, hipStream_t s, bool defer = false);
  int gs_passes = 1;  // classical Gram-Schmidt passes per inner iteration (1 or 2)
  bool defer_finish = true;  // leave one-pass finishes to the next preconditioner launch
  GsFinish pend;             // the finish left by the last Gram-Schmidt (p3 NULL: none)
  void flush_finish(hipStream_t s);  // run a pending finish now
  void norm2_32(const float2* w, double2* out, hipStream_t s);  // out[0] = |w|^2
  // r = b - r, out[0] = |r|^2 (FP64; pinned out and device out_dev)
  void residual(const double2* b, double2* r, double2* out, double2* out_dev, hipStream_t s);
  // out[j] (and out_dev[j]) = (sum over blocks of npart[2 b + j], 0), j < count (<= 2): the
  // squared norms whose block partials the FP64 angular epilogue wrote
  void finish_norms(long long blocks, int count, double2* out, double2* out_dev, hipStream_t s);
  bool ss_one_wg = false;  // S x S with one workgroup per channel (else a block of r1 rows each)
  // out = S x S in (FP32, out-of-place; scratch has n elements)
  void overlap_multiply32(const Repr& r, const float2* in, float2* out, float2* scratch, hipStream_t s);
  std::vector<DeviceBuffer<float>> u32 = std::vector<DeviceBuffer<float>>(kMaxRank);
  // Flexible GMRES(m) on A d = rhs, d0 = 0, right preconditioner P_j (j = iteration), then
  // xout = base + d (FP64; base may alias xout).  rhs = NULL: V_0 already holds rhs / |rhs|
  // and beta = |rhs|, or, with first_read, beta is set by first_read at the first host
  // synchronisation, which may also return false: the solve is then abandoned without
  // touching xout and -1 is returned.  apply_P(j, v, n2, z, wb): with n2 (a device squared
  // norm), wb <- v / sqrt(n2->x), the normalised basis vector V_j, and z = P_j wb; without,
  // z = P_j v (v = V_0, already normalised).  For j >= 1, v is wnext, where the Gram-Schmidt
  // update of the previous iteration left the new direction (out of place, so that the
  // preconditioner never reads what it writes).  The first
  // pre_iters iterations run before the host reads anything; *needed (may be NULL) receives
  // the iteration count at which the tolerance was first met.
  // k iterations with no host read; least squares and combination on the device
  void fgmres32_device(const std::function<void(const float2*, float2*)>& apply_A,
                       const std::function<void(int, const float2*, const double2*, float2*, float2*)>& apply_P, int k,
                       const double2* beta2_dev, const double2* base, double2* xout, int est, hipStream_t s);
  int fgmres32(const std::function<void(const float2*, float2*)>& apply_A,
               const std::function<void(int, const float2*, const double2*, float2*, float2*)>& apply_P, const float2* rhs,
               double rtol, int max_it, double* rel_out, hipStream_t s, double beta, const double2* base,
               double2* xout, const std::function<bool(double&)>& first_read = nullptr, int pre_iters = 1,
               int* needed = nullptr);
};

struct MixedCayley {
  Hamiltonian* ham = nullptr;
  const FHam* fham = nullptr;        // factored multiply for both precisions when set
  const Kron* kron = nullptr;        // FP32 base preconditioner (required)
  Learned* learned = nullptr;        // optional correction of the Kronecker preconditioner
  int learned_mode = 0;              // where it acts: 0 first direction of the first correction,
                                     // 1 first direction of every correction, 2 every direction
  Cayley* fallback = nullptr;        // optional plain FP64 solver (exact block preconditioner)
  int m = 20;                        // inner FGMRES basis
  double tol_rel = 1e-10;            // outer stop: |b - A x| <= tol_rel |b|
  double inner_rtol = 1e-5;          // inner stop: relative to the current residual
  int max_outer = 8;
  int fallback_iterations = 40;      // inner iterations of one step beyond which the plain solver redoes it
  bool overlap_from_solve = true;    // the FP32 S z of a preconditioner output comes from its Kronecker solve
  MixedWork work;
  MixedStats stats;
  // reuse of the last FP64 radial products across steps (factored operator only)
  bool z_valid = false;
  unsigned long long z_gen = 0;
  const double2* z_psi = nullptr;
  std::vector<double> scal;
  std::vector<int> pre_its;  // per correction: inner iterations run before the first host read
  bool device_mode = false;  // after a first host-driven step: one host synchronisation per step
  int prev_corrections = 0;  // corrections the last converged step needed
  // Graph mode (with device mode and a sparse angular form): the device-mode step is
  // captured once per (corrections, inner counts, reuse, psi) and replayed; the time
  // scalings reach the kernels through scal_dev.
  bool graph_mode = false;
  bool graph_ok = true;          // cleared when a capture fails (the step then runs eagerly)
  DeviceBuffer<double> scal_dev;
  double* scal_host = nullptr;   // pinned staging of the scalings
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<long long> gkey, eager_key;  // the captured sequence's key; the last eager one's
  MixedStats gstats;             // the counts the captured sequence adds to a step's statistics
  int graph_captures = 0, graph_launches = 0;
  void step(double2* psi, double dt, const double* scalings, hipStream_t s);
  ~MixedCayley();
};

}  // namespace pp

struct pp_kron {
  pp::Kron k;
};
struct pp_mixed {
  pp::MixedCayley c;
};

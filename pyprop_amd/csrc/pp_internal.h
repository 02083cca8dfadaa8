// Internal types shared between the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pp_common.h"

namespace pp {

constexpr int kMaxTerms = 8;     // tensor potentials fused into one launch
constexpr int kMaxRank = 4;
constexpr int kMaxDiaRows = 512; // innermost extent handled by the banded kernel

struct TPGeometry {
  int rank = 0;
  int ids[kMaxRank] = {0, 0, 0, 0};
  int pot_ext[kMaxRank] = {0, 0, 0, 0};
  int src_ext[kMaxRank] = {0, 0, 0, 0};
  std::vector<int32_t> pairs[kMaxRank];  // flattened (P, 2) for Simp/Herm ranks
};

struct TPPlanData {
  std::vector<int4> tuples;
  std::vector<int> fib_ptr;
  std::vector<int> fib_order;
  long long n_fibres = 0;
  int non_empty = 0;
  int n = 0;
};

struct TPPlanDevice {
  DeviceBuffer<int4> tuples;
  DeviceBuffer<int> fib_ptr;
  DeviceBuffer<int> fib_order;
  long long n_fibres = 0;
  long long tuple_count = 0;
  int non_empty = 0;
  int n = 0;
};

struct TPTerm {
  TPGeometry geom;
  bool is_real = false;
  bool dia = false;
  int inner_kind = 0;
  int bw = 0;
  long long pot_elements = 0;
  long long pot_fibres = 0;
  long long stream_bytes = 0;
  DeviceBuffer<double> dia_real;
  DeviceBuffer<double2> dia_cplx;
  DeviceBuffer<float> dia_real_f;
  DeviceBuffer<float2> dia_cplx_f;
  DeviceBuffer<double2> vref;      // generic path: reference layout copy
  DeviceBuffer<int> row_ptr;
  DeviceBuffer<int4> entries;
  TPPlanDevice plan;               // single-term plan
};

TPPlanData build_plan_host(const std::vector<const TPTerm*>& terms);
void upload_plan(const TPPlanData& pd, TPPlanDevice& dv, hipStream_t s);
bool terms_fusable(const std::vector<const TPTerm*>& terms);
void tp_run_f64(const TPPlanDevice& plan, const std::vector<const TPTerm*>& terms,
                const double2* scales, const double2* src, double2* dst, const double2* addend,
                double2 alpha, bool accumulate, hipStream_t s);
void tp_run_f32(const TPPlanDevice& plan, const std::vector<const TPTerm*>& terms,
                const float2* scales, const float2* src, float2* dst, const float2* addend,
                float2 alpha, bool accumulate, hipStream_t s);
void tp_ensure_f32(const TPTerm& t, hipStream_t s);

// BLAS-1 helpers (blas.hip)
struct Reducer {
  // scratch for two-stage deterministic reductions
  DeviceBuffer<double2> partial;
  DeviceBuffer<double2> result;
  int blocks = 0;
  void ensure(int nblocks, int width);
};
void vec_scale_copy(long long n, const double2* x, double2* y, const double2* scale_dev,
                    bool invert_real, hipStream_t s);
int reduction_blocks(long long n);
// out[0..m-1] = V_i^H w for i < m (V rows of length n, stride ld), out[m] = ||w||^2 (real part)
void block_dot(long long n, int m, const double2* V, long long ld, const double2* w, double2* out_dev,
               Reducer& red, hipStream_t s);
// w -= sum_i V_i h_i ; out_dev[0] = ||w||^2
void block_update(long long n, int m, const double2* V, long long ld, const double2* h_dev,
                  double2* w, double2* norm2_dev, Reducer& red, hipStream_t s);
void vec_dot(long long n, const double2* x, const double2* y, double2* out_dev, Reducer& red, hipStream_t s);
// the same dot with caller-provided reduction scratch (2 reduction_blocks(n) values)
void vec_dot_into(long long n, const double2* x, const double2* y, double2* out_dev, double2* partial,
                  hipStream_t s);
void vec_norm2(long long n, const double2* x, double2* out_dev, Reducer& red, hipStream_t s);
void vec_axpy_dev(long long n, const double2* alpha_dev, bool negate, const double2* x, double2* y, hipStream_t s);
void vec_combine(long long n, double2 a, const double2* x, double2 b, const double2* y, double2* out, hipStream_t s);
void vec_lincomb(long long n, int m, const double2* V, long long ld, const double2* coef_dev,
                 const double2* scale_dev, double2* out, hipStream_t s);
void vec_copy(long long n, const double2* x, double2* y, hipStream_t s);
void vec_zero(long long n, double2* y, hipStream_t s);

// overlap (overlap.hip)
struct OverlapFactor {
  int n = 0;
  int sd = 0;                    // superdiagonals
  std::vector<double> u_host;    // (n, sd+1) upper band, [j][sd + i - j]
  DeviceBuffer<double> u;        // same layout on device
  mutable DeviceBuffer<double2> work;  // out-of-place passes of the multiply ops (grown on demand)
};
// work: out-of-place buffer of the multiply ops (as many elements as data); NULL uses ov.work,
// grown on demand
void overlap_apply(const OverlapFactor& ov, int op, bool conj, int ndim, const int* extents, int rank,
                   double2* data, hipStream_t s, double2* work = nullptr);

}  // namespace pp

struct pp_tensor_potential {
  pp::TPTerm term;
};

struct pp_overlap {
  pp::OverlapFactor f;
};

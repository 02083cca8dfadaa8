// Tensor-potential multiply on MI355X.
//
// Semantics follow the generated Fortran kernels of
// core/tensorpotential/tensorpotentialmultiply_generator.py:
//   dest(row_0..row_R-1) += [conj] V(i_0..i_R-1) * scaling * source(col_0..col_R-1)
// summed over each rank's (row, col, i) enumeration (Simp 302-363, Herm 370-398,
// Diag 403-432, Ident 436-462, Band 467-618, BandNH 622-762, Dense 1227-1330).
//
// Design.  The outer ranks (all but the innermost) only enumerate index tuples; the
// plan lists them on the host once, in the reference loop order, grouped by destination
// fibre (a fibre is the innermost-rank vector at fixed outer indices).  One workgroup
// owns one destination fibre, so no atomics are needed and the summation order is fixed
// (deterministic).  Its four waves split the fibre's tuple list in four contiguous parts
// whose partial sums are added in order at the end.  Workgroups are launched in order of
// decreasing tuple count (longest first) to balance the tail.
//
// Innermost rank:
//  * banded storages (BandNH = zgbmv, Band = zhbmv) are re-laid out once into
//    diagonal-major form D[fibre][d][i] = A(i, i + d - hb): lane i streams row i with
//    fully coalesced loads (consecutive lanes read consecutive addresses), the source
//    window is staged in LDS per wave.  If every imaginary part is zero the diagonals are
//    stored as float64 (8 B/element instead of 16).  This path is HBM-bound.
//  * all other storages go through a generic per-row entry list (CSR) over the
//    reference layout; correct for every storage, not tuned.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <numeric>
#include <vector>

#include "pp_common.h"
#include "pp_internal.h"

namespace pp {

// ----------------------------------------------------------------------------------
// host: per-rank enumerations in reference loop order
// ----------------------------------------------------------------------------------
struct RankEntry {
  int row, col, pidx, toggle;
};

static std::vector<RankEntry> enumerate_rank(int id, int potExt, int srcExt, const int32_t* pairs,
                                             int pairCount) {
  std::vector<RankEntry> e;
  switch (id) {
    case PP_STORAGE_SIMP:
      PP_REQUIRE(pairs && pairCount == potExt, "Simp rank needs pairs with count == potential extent");
      for (int i = 0; i < potExt; ++i) e.push_back({pairs[2 * i], pairs[2 * i + 1], i, 0});
      break;
    case PP_STORAGE_HERM:
      PP_REQUIRE(pairs && pairCount == potExt, "Herm rank needs pairs with count == potential extent");
      for (int i = 0; i < potExt; ++i) {
        e.push_back({pairs[2 * i], pairs[2 * i + 1], i, 0});
        e.push_back({pairs[2 * i + 1], pairs[2 * i], i, 1});
      }
      break;
    case PP_STORAGE_DIAG:
      for (int i = 0; i < potExt; ++i) e.push_back({i, i, i, 0});
      break;
    case PP_STORAGE_IDENT:
      for (int r = 0; r < srcExt; ++r) e.push_back({r, r, 0, 0});
      break;
    case PP_STORAGE_DENSE:
      for (int r = 0; r < srcExt; ++r)
        for (int c = 0; c < srcExt; ++c) e.push_back({r, c, r * srcExt + c, 0});
      break;
    case PP_STORAGE_BAND: {
      PP_REQUIRE(srcExt > 0 && potExt % srcExt == 0, "Band potential extent must be a multiple of N");
      int band = potExt / srcExt;
      for (int r = 0; r < srcExt; ++r)
        for (int c = r; c < std::min(r + band, srcExt); ++c) e.push_back({r, c, r * band + c - r, 0});
      for (int hr = 0; hr < srcExt; ++hr)
        for (int hc = hr + 1; hc < std::min(hr + band, srcExt); ++hc)
          e.push_back({hc, hr, hr * band + hc - hr, 1});
      break;
    }
    case PP_STORAGE_BANDNH: {
      PP_REQUIRE(srcExt > 0 && potExt % srcExt == 0, "BandNH potential extent must be a multiple of N");
      int bw = potExt / srcExt;
      PP_REQUIRE(bw % 2 == 1, "BandNH band count must be odd");
      int b = (bw - 1) / 2;
      for (int c = 0; c < srcExt; ++c)
        for (int r = std::max(0, c - b); r <= std::min(srcExt - 1, c + b); ++r)
          e.push_back({r, c, c * bw + b + r - c, 0});
      break;
    }
    default:
      throw Error(PP_ERR_INVALID, "unknown storage id " + std::to_string(id));
  }
  for (auto& x : e) {
    PP_REQUIRE(x.row >= 0 && x.row < srcExt && x.col >= 0 && x.col < srcExt,
               "pair index out of range for rank extent");
    PP_REQUIRE(x.pidx >= 0 && x.pidx < potExt, "potential index out of range");
  }
  return e;
}

// ----------------------------------------------------------------------------------
// device kernels
// ----------------------------------------------------------------------------------
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ inline T cz();
template <>
__device__ inline double2 cz<double2>() { return make_double2(0.0, 0.0); }
template <>
__device__ inline float2 cz<float2>() { return make_float2(0.f, 0.f); }

// acc += (conj?) v * x for real or complex v
template <class S, class C>
__device__ inline void acc_vx(C& acc, S v, C x, bool) {
  acc.x = fma(v, x.x, acc.x);
  acc.y = fma(v, x.y, acc.y);
}
template <class C>
__device__ inline void acc_vx(C& acc, C v, C x, bool cj) {
  const auto vy = cj ? -v.y : v.y;
  acc.x = fma(v.x, x.x, acc.x);
  acc.x = fma(-vy, x.y, acc.x);
  acc.y = fma(v.x, x.y, acc.y);
  acc.y = fma(vy, x.x, acc.y);
}

template <class C>
__device__ inline void acc_scaled(C& acc, C s, C p) {
  acc.x = fma(s.x, p.x, acc.x);
  acc.x = fma(-s.y, p.y, acc.x);
  acc.y = fma(s.x, p.y, acc.y);
  acc.y = fma(s.y, p.x, acc.y);
}

// streaming (non-temporal) loads: the diagonals are read exactly once per product
typedef double pp_d2v __attribute__((ext_vector_type(2)));
typedef float pp_f2v __attribute__((ext_vector_type(2)));
__device__ inline double ld_stream(const double* p) { return __builtin_nontemporal_load(p); }
__device__ inline float ld_stream(const float* p) { return __builtin_nontemporal_load(p); }
__device__ inline double2 ld_stream(const double2* p) {
  const pp_d2v v = __builtin_nontemporal_load(reinterpret_cast<const pp_d2v*>(p));
  return make_double2(v.x, v.y);
}
__device__ inline float2 ld_stream(const float2* p) {
  const pp_f2v v = __builtin_nontemporal_load(reinterpret_cast<const pp_f2v*>(p));
  return make_float2(v.x, v.y);
}

template <class C>
struct DiaArgs {
  const int4* tuples;   // {src fibre, potential fibre, (term << 1) | conj, 0}
  const int* fib_ptr;   // per destination fibre, tuple range
  const int* fib_order; // launch order
  const void* dia[kMaxTerms];
  C scale[kMaxTerms];
  int n;                // innermost extent
  const C* src;
  C* dst;
  const C* addend;      // optional
  C alpha;
  int accumulate;       // dst_out = (accumulate ? dst_in : 0) + addend + alpha * sum
};

template <class V, class C, int BW, int NR>
__global__ __launch_bounds__(256) void tp_dia_kernel(DiaArgs<C> a) {
  constexpr int HB = (BW - 1) / 2;
  constexpr int WIN = NR * 64 + BW;
  __shared__ C xs[4][WIN];  // per-wave source window; reused for the final reduction
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int fib = a.fib_order[blockIdx.x];
  const int t0 = a.fib_ptr[fib];
  const int cnt = a.fib_ptr[fib + 1] - t0;
  const int w0 = t0 + (int)(((long long)cnt * wave) >> 2);
  const int w1 = t0 + (int)(((long long)cnt * (wave + 1)) >> 2);
  const int n = a.n;
  C* xw = xs[wave];
  for (int j = lane; j < WIN; j += 64) xw[j] = cz<C>();
  C acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = cz<C>();
  wave_sync();
  for (int t = w0; t < w1; ++t) {
    const int4 tp = a.tuples[t];
    const C* x = a.src + (long long)tp.x * n;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = lane + 64 * r;
      if (i < n) xw[HB + i] = x[i];
    }
    wave_sync();
    const int term = tp.z >> 1;
    const bool cj = tp.z & 1;
    const V* v = reinterpret_cast<const V*>(a.dia[term]) + (long long)tp.y * BW * n;
    const C s = a.scale[term];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = lane + 64 * r;
      if (i < n) {
        V vv[BW];
#pragma unroll
        for (int d = 0; d < BW; ++d) vv[d] = ld_stream(v + d * n + i);
        C part = cz<C>();
#pragma unroll
        for (int d = 0; d < BW; ++d) acc_vx(part, vv[d], xw[i + d], cj);
        acc_scaled(acc[r], s, part);
      }
    }
    wave_sync();
  }
  __syncthreads();
  C(*red)[WIN] = xs;
#pragma unroll
  for (int r = 0; r < NR; ++r) red[wave][lane + 64 * r] = acc[r];
  __syncthreads();
  if (wave == 0) {
    C* out = a.dst + (long long)fib * n;
    const C* add = a.addend ? a.addend + (long long)fib * n : nullptr;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = lane + 64 * r;
      if (i < n) {
        C s = red[0][i];
        s.x += red[1][i].x; s.y += red[1][i].y;
        s.x += red[2][i].x; s.y += red[2][i].y;
        s.x += red[3][i].x; s.y += red[3][i].y;
        C o = a.accumulate ? out[i] : cz<C>();
        if (add) { o.x += add[i].x; o.y += add[i].y; }
        acc_scaled(o, a.alpha, s);
        out[i] = o;
      }
    }
  }
}

// generic path: entries per inner row (CSR) over the reference layout
template <class C>
struct GenArgs {
  const int4* tuples;
  const int* fib_ptr;
  const int* fib_order;
  const int* row_ptr;      // n + 1
  const int4* entries;     // {col, potential index, flags, 0}; flags: 1 conj, 2 real part, 4 ignore tuple conj
  const void* vref[kMaxTerms];
  C scale[kMaxTerms];
  long long pin;           // innermost potential extent
  int n;
  const C* src;
  C* dst;
  const C* addend;
  C alpha;
  int accumulate;
};

template <class C>
__global__ __launch_bounds__(64) void tp_generic_kernel(GenArgs<C> a) {
  const int fib = a.fib_order[blockIdx.x];
  const int t0 = a.fib_ptr[fib];
  const int t1 = a.fib_ptr[fib + 1];
  const int n = a.n;
  C* out = a.dst + (long long)fib * n;
  const C* add = a.addend ? a.addend + (long long)fib * n : nullptr;
  for (int i = threadIdx.x; i < n; i += 64) {
    C acc = cz<C>();
    const int e0 = a.row_ptr[i], e1 = a.row_ptr[i + 1];
    for (int t = t0; t < t1; ++t) {
      const int4 tp = a.tuples[t];
      const int term = tp.z >> 1;
      const bool tcj = tp.z & 1;
      const C* v = reinterpret_cast<const C*>(a.vref[term]) + (long long)tp.y * a.pin;
      const C* x = a.src + (long long)tp.x * n;
      C part = cz<C>();
      for (int e = e0; e < e1; ++e) {
        const int4 en = a.entries[e];
        C vv = v[en.y];
        if (en.z & 2) vv.y = 0;
        const bool cj = ((en.z & 4) ? false : tcj) ^ (bool)(en.z & 1);
        acc_vx(part, vv, x[en.x], cj);
      }
      acc_scaled(acc, a.scale[term], part);
    }
    C o = a.accumulate ? out[i] : cz<C>();
    if (add) { o.x += add[i].x; o.y += add[i].y; }
    acc_scaled(o, a.alpha, acc);
    out[i] = o;
  }
}

// repack reference band layouts into diagonal-major form
//   BandNH: A(i, j) = V[j*bw + hb + i - j]
//   Band  : A(i, j) = V[j*k + i - j] (i > j), conj(V[i*k + j - i]) (i < j), Re V[i*k] (i == j)
template <class OutT>
__global__ void repack_band_kernel(const double2* __restrict__ vref, OutT* __restrict__ out,
                                   long long fibres, int n, int bw, int kind, long long pin) {
  const int hb = (bw - 1) / 2;
  const long long total = fibres * bw * (long long)n;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n);
    const int d = (int)((idx / n) % bw);
    const long long f = idx / ((long long)n * bw);
    const int j = i + d - hb;
    double2 val = make_double2(0.0, 0.0);
    if (j >= 0 && j < n) {
      const double2* v = vref + f * pin;
      if (kind == PP_STORAGE_BANDNH) {
        val = v[(long long)j * bw + hb + i - j];
      } else {
        const int k = hb + 1;
        if (i > j) val = v[(long long)j * k + i - j];
        else if (i < j) { val = v[(long long)i * k + j - i]; val.y = -val.y; }
        else { val = v[(long long)i * k]; val.y = 0.0; }
      }
    }
    if constexpr (sizeof(OutT) == sizeof(double)) out[idx] = (OutT)val.x;
    else if constexpr (sizeof(OutT) == sizeof(double2)) out[idx] = val;
    else if constexpr (sizeof(OutT) == sizeof(float)) out[idx] = (OutT)val.x;
    else out[idx] = make_float2((float)val.x, (float)val.y);
  }
}

__global__ void max_abs_imag_kernel(const double2* __restrict__ v, long long n, unsigned long long* flag) {
  bool nz = false;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    nz |= (v[i].y != 0.0);
  if (__any(nz) && (threadIdx.x & 63) == 0) atomicOr(flag, 1ull);
}

__global__ void expand_kernel(int terms, long long p0, long long rc, const double* __restrict__ coeffs,
                              const double* __restrict__ radial, double2* __restrict__ out) {
  const long long total = p0 * rc;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long a = idx / rc, r = idx % rc;
    double v = 0.0;
    for (int t = 0; t < terms; ++t) v = fma(coeffs[t * p0 + a], radial[t * rc + r], v);
    out[idx] = make_double2(v, 0.0);
  }
}

__global__ void convert_f64_f32_kernel(const double* __restrict__ in, float* __restrict__ out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// ----------------------------------------------------------------------------------
// plan construction
// ----------------------------------------------------------------------------------
struct Tuple {
  int dst, src, pot, conj;
};

static std::vector<Tuple> outer_tuples(const TPGeometry& g) {
  const int R = g.rank;
  std::vector<std::vector<RankEntry>> per(R > 1 ? R - 1 : 0);
  for (int r = 0; r + 1 < R; ++r)
    per[r] = enumerate_rank(g.ids[r], g.pot_ext[r], g.src_ext[r], g.pairs[r].empty() ? nullptr : g.pairs[r].data(),
                            (int)(g.pairs[r].size() / 2));
  std::vector<Tuple> out;
  if (R == 1) {
    out.push_back({0, 0, 0, 0});
    return out;
  }
  // strides over the outer dims
  std::vector<long long> sstride(R - 1), pstride(R - 1);
  long long s = 1, p = 1;
  for (int r = R - 2; r >= 0; --r) {
    sstride[r] = s;
    pstride[r] = p;
    s *= g.src_ext[r];
    p *= g.pot_ext[r];
  }
  PP_REQUIRE(s < (1ll << 31) && p < (1ll << 31), "outer fibre count exceeds int32");
  long long total = 1;
  for (auto& v : per) total *= (long long)v.size();
  PP_REQUIRE(total < (1ll << 31), "tuple count exceeds int32");
  out.reserve((size_t)total);
  std::vector<size_t> idx(R - 1, 0);
  // iterate the nested loops (rank 0 outermost)
  while (true) {
    long long d = 0, sc = 0, pp_ = 0;
    int cj = 0;
    for (int r = 0; r < R - 1; ++r) {
      const RankEntry& e = per[r][idx[r]];
      d += e.row * sstride[r];
      sc += e.col * sstride[r];
      pp_ += e.pidx * pstride[r];
      cj ^= e.toggle;
    }
    out.push_back({(int)d, (int)sc, (int)pp_, cj});
    int r = R - 2;
    while (r >= 0) {
      if (++idx[r] < per[r].size()) break;
      idx[r] = 0;
      --r;
    }
    if (r < 0) break;
  }
  return out;
}

static void build_inner_csr(const TPGeometry& g, std::vector<int>& row_ptr, std::vector<int4>& entries) {
  const int R = g.rank;
  const int id = g.ids[R - 1];
  const int n = g.src_ext[R - 1];
  const int pin = g.pot_ext[R - 1];
  std::vector<std::vector<int4>> rows(n);
  if (id == PP_STORAGE_BAND) {
    PP_REQUIRE(pin % n == 0, "Band potential extent must be a multiple of N");
    const int k = pin / n;
    for (int i = 0; i < n; ++i) {
      for (int j = std::max(0, i - k + 1); j < i; ++j) rows[i].push_back(make_int4(j, j * k + i - j, 4, 0));
      rows[i].push_back(make_int4(i, i * k, 2 | 4, 0));
      for (int j = i + 1; j < std::min(n, i + k); ++j) rows[i].push_back(make_int4(j, i * k + j - i, 1 | 4, 0));
    }
  } else {
    auto e = enumerate_rank(id, pin, n, g.pairs[R - 1].empty() ? nullptr : g.pairs[R - 1].data(),
                            (int)(g.pairs[R - 1].size() / 2));
    for (auto& x : e) rows[x.row].push_back(make_int4(x.col, x.pidx, x.toggle, 0));
  }
  row_ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + (int)rows[i].size();
  entries.clear();
  for (auto& r : rows) entries.insert(entries.end(), r.begin(), r.end());
}

TPPlanData build_plan_host(const std::vector<const TPTerm*>& terms) {
  PP_REQUIRE(!terms.empty() && terms.size() <= (size_t)kMaxTerms, "1..8 terms per plan");
  const TPGeometry& g0 = terms[0]->geom;
  TPPlanData pd;
  const int R = g0.rank;
  pd.n = g0.src_ext[R - 1];
  long long nf = 1;
  for (int r = 0; r + 1 < R; ++r) nf *= g0.src_ext[r];
  pd.n_fibres = nf;
  std::vector<std::vector<int4>> byfib(nf);
  for (size_t t = 0; t < terms.size(); ++t) {
    const TPGeometry& g = terms[t]->geom;
    PP_REQUIRE(g.rank == R, "terms must share the rank");
    for (int r = 0; r < R; ++r) PP_REQUIRE(g.src_ext[r] == g0.src_ext[r], "terms must share the source shape");
    const bool ignore_conj = g.ids[R - 1] == PP_STORAGE_BAND;  // zhbmv takes no conjugate flag
    for (const Tuple& tp : outer_tuples(g))
      byfib[tp.dst].push_back(make_int4(tp.src, tp.pot, (int)(t << 1) | (ignore_conj ? 0 : tp.conj), 0));
  }
  pd.fib_ptr.assign(nf + 1, 0);
  for (long long f = 0; f < nf; ++f) pd.fib_ptr[f + 1] = pd.fib_ptr[f] + (int)byfib[f].size();
  pd.tuples.reserve(pd.fib_ptr[nf]);
  for (auto& v : byfib) pd.tuples.insert(pd.tuples.end(), v.begin(), v.end());
  pd.fib_order.resize(nf);
  std::iota(pd.fib_order.begin(), pd.fib_order.end(), 0);
  std::stable_sort(pd.fib_order.begin(), pd.fib_order.end(), [&](int a, int b) {
    return (pd.fib_ptr[a + 1] - pd.fib_ptr[a]) > (pd.fib_ptr[b + 1] - pd.fib_ptr[b]);
  });
  pd.non_empty = 0;
  for (long long f = 0; f < nf; ++f)
    if (pd.fib_ptr[f + 1] > pd.fib_ptr[f]) ++pd.non_empty;
  return pd;
}

void upload_plan(const TPPlanData& pd, TPPlanDevice& dv, hipStream_t s) {
  dv.tuples.upload(pd.tuples.data(), pd.tuples.size(), s);
  dv.fib_ptr.upload(pd.fib_ptr.data(), pd.fib_ptr.size(), s);
  dv.fib_order.upload(pd.fib_order.data(), pd.fib_order.size(), s);
  dv.n = pd.n;
  dv.n_fibres = pd.n_fibres;
  dv.non_empty = pd.non_empty;
  dv.tuple_count = (long long)pd.tuples.size();
}

static bool is_band(int id) { return id == PP_STORAGE_BANDNH || id == PP_STORAGE_BAND; }

void tp_create_impl(int rank, const int* ids, const double2* potential, const int* pot_ext,
                    const int* src_ext, const int32_t* const* pairs, const int* pair_counts,
                    hipStream_t s, TPTerm& term) {
  PP_REQUIRE(rank >= 1 && rank <= kMaxRank, "rank must be 1..4");
  PP_REQUIRE(potential != nullptr, "potential is NULL");
  TPGeometry& g = term.geom;
  g.rank = rank;
  long long pcount = 1;
  for (int r = 0; r < rank; ++r) {
    g.ids[r] = ids[r];
    g.pot_ext[r] = pot_ext[r];
    g.src_ext[r] = src_ext[r];
    PP_REQUIRE(pot_ext[r] > 0 && src_ext[r] > 0, "extents must be positive");
    pcount *= pot_ext[r];
    if (ids[r] == PP_STORAGE_SIMP || ids[r] == PP_STORAGE_HERM) {
      PP_REQUIRE(pairs && pairs[r] && pair_counts && pair_counts[r] == pot_ext[r],
                 "Simp/Herm rank needs a pair list of the potential extent");
      g.pairs[r].assign(pairs[r], pairs[r] + 2 * (size_t)pair_counts[r]);
    } else {
      g.pairs[r].clear();
    }
    if (ids[r] == PP_STORAGE_DIAG) PP_REQUIRE(pot_ext[r] <= src_ext[r], "Diag extent exceeds source extent");
    if (ids[r] == PP_STORAGE_IDENT) PP_REQUIRE(pot_ext[r] == 1, "Ident potential extent must be 1");
    if (ids[r] == PP_STORAGE_DENSE) PP_REQUIRE(pot_ext[r] == src_ext[r] * src_ext[r], "Dense potential extent must be N*N");
  }
  const int inner = ids[rank - 1];
  const int n = src_ext[rank - 1];
  const long long pin = pot_ext[rank - 1];
  const long long pfib = pcount / pin;
  term.pot_elements = pcount;
  term.pot_fibres = pfib;

  // realness of the potential
  {
    DeviceBuffer<unsigned long long> flag(1);
    PP_HIP(hipMemsetAsync(flag.ptr, 0, sizeof(unsigned long long), s));
    int blocks = (int)std::min<long long>(4096, (pcount + 255) / 256);
    hipLaunchKernelGGL(max_abs_imag_kernel, dim3(blocks), dim3(256), 0, s, potential, pcount, flag.ptr);
    PP_HIP(hipGetLastError());
    unsigned long long h = 0;
    PP_HIP(hipMemcpyAsync(&h, flag.ptr, sizeof(h), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    term.is_real = (h == 0);
  }

  term.inner_kind = inner;
  if (is_band(inner) && n <= kMaxDiaRows) {
    PP_REQUIRE(pin % n == 0, "band potential extent must be a multiple of the source extent");
    int bw = (inner == PP_STORAGE_BANDNH) ? (int)(pin / n) : (int)(2 * (pin / n) - 1);
    PP_REQUIRE(bw % 2 == 1 && bw >= 1, "band width must be odd");
    term.dia = true;
    term.bw = bw;
    const long long total = pfib * bw * (long long)n;
    int blocks = (int)std::min<long long>(65536, (total + 255) / 256);
    if (term.is_real) {
      term.dia_real.resize(total);
      hipLaunchKernelGGL(repack_band_kernel<double>, dim3(blocks), dim3(256), 0, s, potential,
                         term.dia_real.ptr, pfib, n, bw, inner, pin);
      term.stream_bytes = total * (long long)sizeof(double);
    } else {
      term.dia_cplx.resize(total);
      hipLaunchKernelGGL(repack_band_kernel<double2>, dim3(blocks), dim3(256), 0, s, potential,
                         term.dia_cplx.ptr, pfib, n, bw, inner, pin);
      term.stream_bytes = total * (long long)sizeof(double2);
    }
    PP_HIP(hipGetLastError());
  } else {
    term.dia = false;
    term.vref.resize(pcount);
    PP_HIP(hipMemcpyAsync(term.vref.ptr, potential, pcount * sizeof(double2), hipMemcpyDeviceToDevice, s));
    std::vector<int> row_ptr;
    std::vector<int4> entries;
    build_inner_csr(g, row_ptr, entries);
    term.row_ptr.upload(row_ptr.data(), row_ptr.size(), s);
    term.entries.upload(entries.data(), entries.size(), s);
    term.stream_bytes = pcount * (long long)sizeof(double2);
  }
  TPPlanData pd = build_plan_host({&term});
  upload_plan(pd, term.plan, s);
  PP_HIP(hipStreamSynchronize(s));
}

void tp_ensure_f32(const TPTerm& t, hipStream_t s) {
  if (!t.dia) throw Error(PP_ERR_INVALID, "FP32 multiply needs a banded innermost storage");
  TPTerm& m = const_cast<TPTerm&>(t);
  const long long total = t.pot_fibres * t.bw * (long long)t.plan.n;
  if (t.is_real) {
    if (m.dia_real_f.count == (size_t)total) return;
    m.dia_real_f.resize(total);
    int blocks = (int)std::min<long long>(65536, (total + 255) / 256);
    hipLaunchKernelGGL(convert_f64_f32_kernel, dim3(blocks), dim3(256), 0, s, t.dia_real.ptr,
                       m.dia_real_f.ptr, total);
  } else {
    if (m.dia_cplx_f.count == (size_t)total) return;
    m.dia_cplx_f.resize(total);
    int blocks = (int)std::min<long long>(65536, (2 * total + 255) / 256);
    hipLaunchKernelGGL(convert_f64_f32_kernel, dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<const double*>(t.dia_cplx.ptr),
                       reinterpret_cast<float*>(m.dia_cplx_f.ptr), 2 * total);
  }
  PP_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------------
// launch
// ----------------------------------------------------------------------------------
template <class V, class C, int BW>
static void launch_dia_bw(const DiaArgs<C>& a, int nr, int blocks, hipStream_t s) {
  // rows per lane rounded up to 1, 2, 4 or 8 (the kernel masks rows >= n)
  if (nr <= 1) hipLaunchKernelGGL((tp_dia_kernel<V, C, BW, 1>), dim3(blocks), dim3(256), 0, s, a);
  else if (nr <= 2) hipLaunchKernelGGL((tp_dia_kernel<V, C, BW, 2>), dim3(blocks), dim3(256), 0, s, a);
  else if (nr <= 4) hipLaunchKernelGGL((tp_dia_kernel<V, C, BW, 4>), dim3(blocks), dim3(256), 0, s, a);
  else if (nr <= 8) hipLaunchKernelGGL((tp_dia_kernel<V, C, BW, 8>), dim3(blocks), dim3(256), 0, s, a);
  else throw Error(PP_ERR_INVALID, "innermost extent too large for the banded kernel");
}

template <class V, class C>
static void launch_dia(const DiaArgs<C>& a, int bw, int nr, int blocks, hipStream_t s) {
  switch (bw) {
    case 1: launch_dia_bw<V, C, 1>(a, nr, blocks, s); break;
    case 3: launch_dia_bw<V, C, 3>(a, nr, blocks, s); break;
    case 5: launch_dia_bw<V, C, 5>(a, nr, blocks, s); break;
    case 7: launch_dia_bw<V, C, 7>(a, nr, blocks, s); break;
    case 9: launch_dia_bw<V, C, 9>(a, nr, blocks, s); break;
    case 11: launch_dia_bw<V, C, 11>(a, nr, blocks, s); break;
    case 13: launch_dia_bw<V, C, 13>(a, nr, blocks, s); break;
    case 15: launch_dia_bw<V, C, 15>(a, nr, blocks, s); break;
    case 17: launch_dia_bw<V, C, 17>(a, nr, blocks, s); break;
    case 19: launch_dia_bw<V, C, 19>(a, nr, blocks, s); break;
    case 21: launch_dia_bw<V, C, 21>(a, nr, blocks, s); break;
    default: throw Error(PP_ERR_INVALID, "band width not instantiated (odd 1..21 supported)");
  }
}

bool terms_fusable(const std::vector<const TPTerm*>& terms) {
  if (terms.empty() || terms.size() > (size_t)kMaxTerms) return false;
  const TPTerm* t0 = terms[0];
  for (auto* t : terms) {
    if (t->dia != t0->dia || t->geom.rank != t0->geom.rank) return false;
    for (int r = 0; r < t0->geom.rank; ++r)
      if (t->geom.src_ext[r] != t0->geom.src_ext[r]) return false;
    if (t->dia && (t->bw != t0->bw || t->is_real != t0->is_real)) return false;
    if (!t->dia && t->geom.pot_ext[t->geom.rank - 1] != t0->geom.pot_ext[t0->geom.rank - 1]) return false;
    if (!t->dia && (t->geom.ids[t->geom.rank - 1] != t0->geom.ids[t0->geom.rank - 1])) return false;
    if (!t->dia && t->geom.pairs[t->geom.rank - 1] != t0->geom.pairs[t0->geom.rank - 1]) return false;
  }
  return true;
}

void tp_run_f64(const TPPlanDevice& plan, const std::vector<const TPTerm*>& terms,
                const double2* scales, const double2* src, double2* dst, const double2* addend,
                double2 alpha, bool accumulate, hipStream_t s) {
  const TPTerm* t0 = terms[0];
  const int blocks = (accumulate && !addend) ? plan.non_empty : (int)plan.n_fibres;
  if (blocks == 0) return;
  if (t0->dia) {
    DiaArgs<double2> a{};
    a.tuples = plan.tuples.ptr;
    a.fib_ptr = plan.fib_ptr.ptr;
    a.fib_order = plan.fib_order.ptr;
    for (size_t i = 0; i < terms.size(); ++i) {
      a.dia[i] = t0->is_real ? (const void*)terms[i]->dia_real.ptr : (const void*)terms[i]->dia_cplx.ptr;
      a.scale[i] = scales[i];
    }
    a.n = plan.n;
    a.src = src;
    a.dst = dst;
    a.addend = addend;
    a.alpha = alpha;
    a.accumulate = accumulate ? 1 : 0;
    const int nr = (plan.n + 63) / 64;
    if (t0->is_real) launch_dia<double, double2>(a, t0->bw, nr, blocks, s);
    else launch_dia<double2, double2>(a, t0->bw, nr, blocks, s);
  } else {
    GenArgs<double2> a{};
    a.tuples = plan.tuples.ptr;
    a.fib_ptr = plan.fib_ptr.ptr;
    a.fib_order = plan.fib_order.ptr;
    a.row_ptr = t0->row_ptr.ptr;
    a.entries = t0->entries.ptr;
    for (size_t i = 0; i < terms.size(); ++i) {
      a.vref[i] = terms[i]->vref.ptr;
      a.scale[i] = scales[i];
    }
    a.pin = t0->geom.pot_ext[t0->geom.rank - 1];
    a.n = plan.n;
    a.src = src;
    a.dst = dst;
    a.addend = addend;
    a.alpha = alpha;
    a.accumulate = accumulate ? 1 : 0;
    hipLaunchKernelGGL(tp_generic_kernel<double2>, dim3(blocks), dim3(64), 0, s, a);
  }
  PP_HIP(hipGetLastError());
}

void tp_run_f32(const TPPlanDevice& plan, const std::vector<const TPTerm*>& terms,
                const float2* scales, const float2* src, float2* dst, const float2* addend,
                float2 alpha, bool accumulate, hipStream_t s) {
  const TPTerm* t0 = terms[0];
  PP_REQUIRE(t0->dia, "FP32 multiply needs a banded innermost storage");
  for (auto* t : terms) tp_ensure_f32(*t, s);
  const int blocks = (accumulate && !addend) ? plan.non_empty : (int)plan.n_fibres;
  if (blocks == 0) return;
  DiaArgs<float2> a{};
  a.tuples = plan.tuples.ptr;
  a.fib_ptr = plan.fib_ptr.ptr;
  a.fib_order = plan.fib_order.ptr;
  for (size_t i = 0; i < terms.size(); ++i) {
    a.dia[i] = t0->is_real ? (const void*)terms[i]->dia_real_f.ptr : (const void*)terms[i]->dia_cplx_f.ptr;
    a.scale[i] = scales[i];
  }
  a.n = plan.n;
  a.src = src;
  a.dst = dst;
  a.addend = addend;
  a.alpha = alpha;
  a.accumulate = accumulate ? 1 : 0;
  const int nr = (plan.n + 63) / 64;
  if (t0->is_real) launch_dia<float, float2>(a, t0->bw, nr, blocks, s);
  else launch_dia<float2, float2>(a, t0->bw, nr, blocks, s);
  PP_HIP(hipGetLastError());
}

}  // namespace pp

// ----------------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------------
using namespace pp;

extern "C" int pp_tp_create(int rank, const int* storage_ids, const pp_cplx* potential,
                            const int* potential_extents, const int* source_extents,
                            const int32_t* const* pairs, const int* pair_counts, pp_stream stream,
                            pp_tensor_potential** out) {
  return guarded([&] {
    PP_REQUIRE(out && storage_ids && potential_extents && source_extents, "NULL argument");
    auto* t = new pp_tensor_potential();
    try {
      tp_create_impl(rank, storage_ids, reinterpret_cast<const double2*>(potential), potential_extents,
                     source_extents, pairs, pair_counts, (hipStream_t)stream, t->term);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

extern "C" int pp_tp_multiply(const pp_tensor_potential* tp, double scaling, const pp_cplx* source,
                              pp_cplx* dest, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(tp && source && dest, "NULL argument");
    double2 sc = make_double2(scaling, 0.0);
    tp_run_f64(tp->term.plan, {&tp->term}, &sc, reinterpret_cast<const double2*>(source),
               reinterpret_cast<double2*>(dest), nullptr, make_double2(1.0, 0.0), true, (hipStream_t)stream);
  });
}

extern "C" int pp_tp_multiply_f32(const pp_tensor_potential* tp, float scaling, const pp_cplxf* source,
                                  pp_cplxf* dest, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(tp && source && dest, "NULL argument");
    float2 sc = make_float2(scaling, 0.f);
    tp_run_f32(tp->term.plan, {&tp->term}, &sc, reinterpret_cast<const float2*>(source),
               reinterpret_cast<float2*>(dest), nullptr, make_float2(1.f, 0.f), true, (hipStream_t)stream);
  });
}

extern "C" int pp_tp_info(const pp_tensor_potential* tp, int64_t* stream_bytes, int64_t* tuple_count,
                          int* is_real, int* inner_kind) {
  return guarded([&] {
    PP_REQUIRE(tp, "NULL plan");
    if (stream_bytes) *stream_bytes = tp->term.stream_bytes;
    if (tuple_count) *tuple_count = tp->term.plan.tuple_count;
    if (is_real) *is_real = tp->term.is_real ? 1 : 0;
    if (inner_kind) *inner_kind = tp->term.dia ? tp->term.inner_kind : -1 - tp->term.inner_kind;
  });
}

extern "C" int pp_tp_destroy(pp_tensor_potential* tp) {
  return guarded([&] { delete tp; });
}

extern "C" int pp_tensor_potential_multiply(int rank, const int* storage_ids, const pp_cplx* potential,
                                            const int* potential_extents, double scaling,
                                            const pp_cplx* source, const int* source_extents,
                                            pp_cplx* dest, const int* dest_extents,
                                            const int32_t* const* pairs, const int* pair_counts,
                                            pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(dest_extents && source_extents, "NULL extents");
    for (int r = 0; r < rank; ++r)
      PP_REQUIRE(dest_extents[r] == source_extents[r], "dest and source extents must match");
    pp_tensor_potential t;
    tp_create_impl(rank, storage_ids, reinterpret_cast<const double2*>(potential), potential_extents,
                   source_extents, pairs, pair_counts, (hipStream_t)stream, t.term);
    double2 sc = make_double2(scaling, 0.0);
    tp_run_f64(t.term.plan, {&t.term}, &sc, reinterpret_cast<const double2*>(source),
               reinterpret_cast<double2*>(dest), nullptr, make_double2(1.0, 0.0), true, (hipStream_t)stream);
    PP_HIP(hipStreamSynchronize((hipStream_t)stream));
  });
}

extern "C" int pp_expand_potential(int terms, int64_t p0, int64_t r_count, const double* coeffs, const double* radial,
                                   pp_cplx* out, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(coeffs && radial && out && terms > 0 && p0 > 0 && r_count > 0, "invalid expand arguments");
    const long long total = p0 * r_count;
    const int blocks = (int)std::min<long long>(65536, (total + 255) / 256);
    hipLaunchKernelGGL(expand_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, terms, (long long)p0,
                       (long long)r_count, coeffs, radial, reinterpret_cast<double2*>(out));
    PP_HIP(hipGetLastError());
  });
}

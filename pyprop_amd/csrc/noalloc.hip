// The S-metric inner product without any device allocation, for callers that cannot
// allocate (code under stream capture, tight serving loops, harnesses that forbid hipMalloc).
//
// pp_repr_inner_product allocates its reduction partials on every call and grows each
// overlap factor's out-of-place pass buffer on demand.  pp_repr_inner_product_noalloc runs
// the same kernels (the copy of b, the S passes of every non-orthogonal rank, the two-stage
// deterministic dot), so its result is bitwise the same, but every intermediate lives in
// one file-scope device array, `scratch`:
//
//   [0, kMaxElements)                 copy of b, multiplied by S in place
//   [kMaxElements, 2 kMaxElements)    out-of-place buffer of the S passes
//   [2 kMaxElements, + kPartials)     reduction partials (2 per block, at most 512 blocks)
//   [.., + 1)                         the reduced value, copied to the host
//
// kMaxElements = 2^20 complex128 elements (16 MiB); the array is 32 MiB + 16 KiB per device
// and is held by every device the library runs on until the process ends.  Larger
// wavefunctions are rejected with PP_ERR_INVALID.  Every call on one device uses that device's
// array: calls that overlap in time on one device (two host threads, two streams) race and
// must be serialised by the caller.  Nothing in the array needs clearing: each call
// overwrites every slot it reads before reading it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pp_common.h"
#include "pp_internal.h"
#include "pp_krylov.h"

namespace {

constexpr long long kMaxElements = PP_NOALLOC_MAX_ELEMENTS;
constexpr long long kPartials = 2 * 512;  // reduction_blocks() is at most 512, 2 values each
constexpr long long kScratchCount = 2 * kMaxElements + kPartials + 1;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double2 scratch[kScratchCount];

std::mutex g_mutex;
double2* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once per device)
double2* scratch_on_current_device() {
  int dev = 0;
  PP_HIP(hipGetDevice(&dev));
  PP_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw pp::Error(PP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double2*>(ptr);
  }
  return g_scratch[dev];
}

}  // namespace

using namespace pp;

extern "C" int pp_repr_inner_product_noalloc(const pp_repr* repr, const pp_cplx* a, const pp_cplx* b,
                                             pp_cplx* result, pp_stream stream) {
  return guarded([&] {
    PP_REQUIRE(repr && a && b && result, "NULL argument");
    const long long n = repr->r.size();
    if (n > kMaxElements)
      throw Error(PP_ERR_INVALID, "pp_repr_inner_product_noalloc: " + std::to_string(n) +
                                  " elements exceed the fixed scratch of " + std::to_string(kMaxElements));
    PP_REQUIRE(reduction_blocks(n) * 2 <= kPartials, "reduction scratch too small");
    double2* base = scratch_on_current_device();
    double2* w = base;
    double2* pass = base + kMaxElements;
    double2* partial = base + 2 * kMaxElements;
    double2* out = partial + kPartials;
    hipStream_t s = (hipStream_t)stream;
    PP_HIP(hipMemcpyAsync(w, b, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    repr->r.overlap(PP_OVERLAP_MULTIPLY, false, w, s, pass);
    vec_dot_into(n, reinterpret_cast<const double2*>(a), w, out, partial, s);
    double2 h;
    PP_HIP(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    PP_HIP(hipStreamSynchronize(s));
    result->re = h.x;
    result->im = h.y;
  });
}

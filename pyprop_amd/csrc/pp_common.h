// Shared helpers for the MI355X (gfx950) propagation library: error handling,
// complex double arithmetic on double2, launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pyprop_amd.h"

namespace pp {

// ---------------------------------------------------------------------------
// errors: every C-ABI entry point catches pp::Error and returns a status code;
// the message is kept per thread for pp_last_error().
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define PP_HIP(expr)                                                              \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess)                                                         \
      throw ::pp::Error(PP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define PP_REQUIRE(cond, msg)                                        \
  do {                                                               \
    if (!(cond)) throw ::pp::Error(PP_ERR_INVALID, std::string(msg)); \
  } while (0)

template <class F>
int guarded(F&& f) {
  try {
    f();
    return PP_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return PP_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return PP_ERR_INTERNAL;
  }
}

// ---------------------------------------------------------------------------
// device memory owned by library objects
// ---------------------------------------------------------------------------
template <class T>
struct DeviceBuffer {
  T* ptr = nullptr;
  size_t count = 0;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void resize(size_t n) {
    release();
    if (n) {
      hipError_t e = hipMalloc(&ptr, n * sizeof(T));
      if (e != hipSuccess)
        throw Error(PP_ERR_ALLOC, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    }
    count = n;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void upload(const T* host, size_t n, hipStream_t s = nullptr) {
    if (n > count) resize(n);
    if (n) PP_HIP(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  size_t bytes() const { return count * sizeof(T); }
};

// ---------------------------------------------------------------------------
// complex double on double2 (re = x, im = y)
// ---------------------------------------------------------------------------
__host__ __device__ inline double2 cmake(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
// acc += a * b
__host__ __device__ inline void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a) * b
__host__ __device__ inline void cfmac(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
__host__ __device__ inline double2 cdiv(double2 a, double2 b) {
  double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

inline double2 to_d2(pp_cplx z) { return make_double2(z.re, z.im); }

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace pp

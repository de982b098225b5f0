// Shared host-side plumbing of libpycd_hip.so: context, errors, buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pycd_hip.h"

struct pycd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  int cu_count = 0;
  double wall_clock_hz = 1e8;  // rate of the kernels' wall_clock64()
  std::vector<uint32_t> simd_keys;  // hardware keys of the SIMDs (probed once)
  // device address of the static scratch of pycd_kmc_advance_static
  double* static_scratch = nullptr;
};

namespace pycd {

// Device address of the plain hop loop's static scratch on `device` (the
// current device), resolved once per device (kmc_plain.hip).
int kmc_static_scratch(int device, double** out);

inline int fail(pycd_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->error = buf;
  return code;
}

#define PYCD_HIP(ctx, expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::pycd::fail((ctx), PYCD_ERR_HIP, "%s failed: %s (%s:%d)", #expr, \
                          hipGetErrorString(e_), __FILE__, __LINE__);         \
  } while (0)

// True when `p` is device (or managed) memory visible to the current device.
inline bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// Owned device allocation.
template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  hipError_t alloc(size_t n) {
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T));
    if (e == hipSuccess) count = n;
    return e;
  }
};

// Copies `n` elements from a host-or-device pointer into device memory.
template <typename T>
hipError_t upload(T* dst, const T* src, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipMemcpyKind kind =
      is_device_ptr(src) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  return hipMemcpyAsync(dst, src, n * sizeof(T), kind, st);
}

template <typename T>
hipError_t download(T* dst, const T* src, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipMemcpyKind kind =
      is_device_ptr(dst) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  return hipMemcpyAsync(dst, src, n * sizeof(T), kind, st);
}

// e**y with e = 2.718281828459045 (the double numpy's `np.e ** y` raises,
// PyCD/core.py:2047).  That double lies below Euler's number, so
// ln(e_double) = 1 - kLnEDefect and e_double**y = exp(y) exp(-y kLnEDefect)
// = exp(y) (1 - y kLnEDefect) to far below an ulp for any rate exponent
// (|y| < 1e6).  One exp and one FMA instead of an FP64 pow; measured
// against the exactly rounded power, within 3.3e-16 relative for
// y in [-60, 5] (the plain exp alone is off by up to 3.3e-15 there).
constexpr double kLnEDefect = 5.318237706605891e-17;

__host__ __device__ inline double numpy_e_pow(double y) {
  const double e = exp(y);
  return fma(-(y * kLnEDefect), e, e);
}

// Scoped device selection.
struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

}  // namespace pycd

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
  // table kernel variant per PYCD_PREC_* (0: the default choice)
  int table_variant[3] = {0, 0, 0};
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

// e_double**y by Tang's table-driven method (P. T. P. Tang, ACM TOMS 15(2),
// 1989): y = (64 m + j) ln2 / 64 + r with |r| <= ln2 / 128, so
// e**y = 2^m 2^(j/64) e^r, with e^r - 1 by its Taylor polynomial to r^6
// (truncation below 4e-20 relative).  The reduction splits ln2 / 64 after
// Cody and Waite into a head of 36 bits, exact times any |64 m + j| < 2^17,
// and a tail; ln(e_double) = 1 - kLnEDefect enters r, whose ulp is far
// below y kLnEDefect.  Within 1.003 ulp of the exactly rounded e_double**y
// for |y| <= 700 (12 million random arguments, against a long-double
// power); 0 below about -745, inf above 709, NaN for NaN.  `T` holds
// 2^(j/64), j = 0..63, exactly rounded (kExp2J64: the bit patterns of the
// doubles nearest to 2^(j/64), from a 60-digit evaluation).
constexpr uint64_t kExp2J64[64] = {
    0x3ff0000000000000, 0x3ff02c9a3e778061, 0x3ff059b0d3158574, 0x3ff0874518759bc8,
    0x3ff0b5586cf9890f, 0x3ff0e3ec32d3d1a2, 0x3ff11301d0125b51, 0x3ff1429aaea92de0,
    0x3ff172b83c7d517b, 0x3ff1a35beb6fcb75, 0x3ff1d4873168b9aa, 0x3ff2063b88628cd6,
    0x3ff2387a6e756238, 0x3ff26b4565e27cdd, 0x3ff29e9df51fdee1, 0x3ff2d285a6e4030b,
    0x3ff306fe0a31b715, 0x3ff33c08b26416ff, 0x3ff371a7373aa9cb, 0x3ff3a7db34e59ff7,
    0x3ff3dea64c123422, 0x3ff4160a21f72e2a, 0x3ff44e086061892d, 0x3ff486a2b5c13cd0,
    0x3ff4bfdad5362a27, 0x3ff4f9b2769d2ca7, 0x3ff5342b569d4f82, 0x3ff56f4736b527da,
    0x3ff5ab07dd485429, 0x3ff5e76f15ad2148, 0x3ff6247eb03a5585, 0x3ff6623882552225,
    0x3ff6a09e667f3bcd, 0x3ff6dfb23c651a2f, 0x3ff71f75e8ec5f74, 0x3ff75feb564267c9,
    0x3ff7a11473eb0187, 0x3ff7e2f336cf4e62, 0x3ff82589994cce13, 0x3ff868d99b4492ed,
    0x3ff8ace5422aa0db, 0x3ff8f1ae99157736, 0x3ff93737b0cdc5e5, 0x3ff97d829fde4e50,
    0x3ff9c49182a3f090, 0x3ffa0c667b5de565, 0x3ffa5503b23e255d, 0x3ffa9e6b5579fdbf,
    0x3ffae89f995ad3ad, 0x3ffb33a2b84f15fb, 0x3ffb7f76f2fb5e47, 0x3ffbcc1e904bc1d2,
    0x3ffc199bdd85529c, 0x3ffc67f12e57d14b, 0x3ffcb720dcef9069, 0x3ffd072d4a07897c,
    0x3ffd5818dcfba487, 0x3ffda9e603db3285, 0x3ffdfc97337b9b5f, 0x3ffe502ee78b3ff6,
    0x3ffea4afa2a490da, 0x3ffefa1bee615a27, 0x3fff50765b6e4540, 0x3fffa7c1819e90d8};
constexpr double kExpInvL = 92.33248261689366;  // 64 / ln2
constexpr double kExpLhi = __builtin_bit_cast(double, 0x3f862e42fefa0000ull);
constexpr double kExpLlo = __builtin_bit_cast(double, 0x3d1cf79abc9e3b3aull);

__device__ inline double e_pow_tab(double y, const double* __restrict__ T) {
#pragma clang fp contract(off)
  const double kf = rint(y * kExpInvL);
  double r = fma(-kf, kExpLhi, y);
  r = fma(-kf, kExpLlo, r);
  r = fma(-y, kLnEDefect, r);
  const int k = int(kf);
  const double t = T[k & 63];
  double p = fma(1.0 / 720, r, 1.0 / 120);
  p = fma(p, r, 1.0 / 24);
  p = fma(p, r, 1.0 / 6);
  p = fma(p, r, 0.5);
  p = fma(p, r * r, r);
  return ldexp(fma(t, p, t), k >> 6);
}

// The reference's hop rate k = nu e**(-dGs / kT) with
// dGs = (lam + dG0)^2 / (4 lam) - V_ab - 0.5 F.d (PyCD/core.py:2040-2047),
// given i4l = 1 / (4 lam), vs = V_ab + 0.5 F.d and nkt = -1 / kT, each
// formed once and rounded once: the exponent moves by a few ulp, the rate
// by |exponent| 1e-16 relative at most.
__device__ inline double marcus_rate(double g0, double lam, double i4l,
                                     double vs, double nkt, double nu,
                                     const double* __restrict__ T) {
#pragma clang fp contract(off)
  const double t = lam + g0;
  const double gs = t * t * i4l - vs;
  return nu * e_pow_tab(gs * nkt, T);
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

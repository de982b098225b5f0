// Wavefront-wide reductions and scans of FP64 values on gfx950: DPP row
// moves, row broadcasts and the permlane16/32 swaps, no LDS.  Every lane of
// the wavefront must be active.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pycd {

// Cross-lane moves through DPP (rows of 16 lanes): a lane whose source is
// outside its row (bound_ctrl) or whose row is masked off reads 0.
template <int kCtrl, int kRowMask, bool kBound>
__device__ inline double dpp_f64(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), kCtrl, kRowMask, 0xf,
                                             kBound);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), kCtrl, kRowMask,
                                             0xf, kBound);
  return __longlong_as_double((int64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

constexpr int kDppRowShr = 0x110;     // + n: lane i reads lane i - n of its row
constexpr int kDppWaveShr1 = 0x138;   // lane i reads lane i - 1
constexpr int kDppRowBcast15 = 0x142; // lane 15 of a row to the next row
constexpr int kDppRowBcast31 = 0x143; // lane 31 to rows 2 and 3

// Inclusive prefix sum over the 64 lanes: Hillis-Steele within each row of
// 16, then the row totals carried across rows.
__device__ inline double wave_incl_scan(double v) {
  v += dpp_f64<kDppRowShr + 1, 0xf, true>(v);
  v += dpp_f64<kDppRowShr + 2, 0xf, true>(v);
  v += dpp_f64<kDppRowShr + 4, 0xf, true>(v);
  v += dpp_f64<kDppRowShr + 8, 0xf, true>(v);
  v += dpp_f64<kDppRowBcast15, 0xa, false>(v);
  v += dpp_f64<kDppRowBcast31, 0xc, false>(v);
  return v;
}

__device__ inline double read_lane(double v, int lane) {
  const int64_t b = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane(int(b), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
  return __longlong_as_double((int64_t(hi) << 32) | lo);
}

// Sum over the wavefront, returned in a scalar register.
__device__ inline double wave_sum(double v) {
  return read_lane(wave_incl_scan(v), 63);
}

constexpr int kDppRowRor = 0x120;  // + n: rotate each row of 16 by n lanes

__device__ inline uint32_t lo32(double x) {
  return uint32_t(__double_as_longlong(x));
}
__device__ inline uint32_t hi32(double x) {
  return uint32_t(uint64_t(__double_as_longlong(x)) >> 32);
}
__device__ inline double from32(uint32_t lo, uint32_t hi) {
  return __longlong_as_double(int64_t((uint64_t(hi) << 32) | lo));
}

// a + b after rows 2-3 of a are exchanged with rows 0-1 of b: rows 0-1
// hold 32 partial sums of a, rows 2-3 32 partial sums of b.
__device__ inline double fold32(double a, double b) {
  const auto l = __builtin_amdgcn_permlane32_swap(lo32(a), lo32(b), false,
                                                  false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi32(a), hi32(b), false,
                                                  false);
  return from32(l[0], h[0]) + from32(l[1], h[1]);
}

// a + b after the odd rows of a are exchanged with the even rows of b:
// rows 0 and 2 hold partial sums of a, rows 1 and 3 of b.
__device__ inline double fold16(double a, double b) {
  const auto l = __builtin_amdgcn_permlane16_swap(lo32(a), lo32(b), false,
                                                  false);
  const auto h = __builtin_amdgcn_permlane16_swap(hi32(a), hi32(b), false,
                                                  false);
  return from32(l[0], h[0]) + from32(l[1], h[1]);
}

// Four sums over the wavefront at once, returned in scalar registers: two
// folds leave 16 partial sums of v[r] in row r, a rotation butterfly sums
// each row.  (Four wave_sum calls would take four full reductions.)
__device__ inline void wave_sum4(const double* v, double* out) {
  double x = fold16(fold32(v[0], v[2]), fold32(v[1], v[3]));
  // a rotation reads a lane of its own row for every lane, so the bound
  // control (which would zero a missing source) never applies; setting it
  // spares the zeroed old value of each move
  x += dpp_f64<kDppRowRor + 8, 0xf, true>(x);
  x += dpp_f64<kDppRowRor + 4, 0xf, true>(x);
  x += dpp_f64<kDppRowRor + 2, 0xf, true>(x);
  x += dpp_f64<kDppRowRor + 1, 0xf, true>(x);
#pragma unroll
  for (int r = 0; r < 4; ++r) out[r] = read_lane(x, 16 * r);
}

__device__ inline double wave_excl_scan(double v) {
  return dpp_f64<kDppWaveShr1, 0xf, true>(wave_incl_scan(v));
}

}  // namespace pycd

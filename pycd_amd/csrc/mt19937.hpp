// CPython's MT19937 (Modules/_randommodule.c semantics) for one wavefront.
//
// The 624-word state lives in LDS; `mti` is the position (0..624) exactly as
// random.getstate()[1][624].  mt_next() returns the next tempered 32-bit
// word, regenerating the block with all 64 lanes when the position reaches
// 624; mt_uniform(a, b) is random.random():
//   ((a >> 5) * 67108864 + (b >> 6)) / 9007199254740992.
// All lanes call these with identical arguments (uniform control flow).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pycd {

constexpr int kMtN = 624;
constexpr int kMtM = 397;

__device__ inline uint32_t mt_mix(uint32_t hi, uint32_t lo, uint32_t far) {
  const uint32_t y = (hi & 0x80000000u) | (lo & 0x7fffffffu);
  return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// Regenerates the block in place with the sequential recurrence's
// dependencies: words [0,227) use old words only, [227,454) use new
// [0,227), [454,623) use new [227,396), word 623 uses new words 0 and 396.
// Within a 64-word round every lane reads before any lane writes, and a
// round never reads a word that the same or an earlier round of its phase
// has written except the new words the recurrence asks for.
__device__ inline void mt_twist(uint32_t* mt, int lane) {
  for (int b = 0; b < kMtN - kMtM; b += 64) {  // [0, 227)
    const int i = b + lane;
    uint32_t v = 0;
    if (i < kMtN - kMtM) v = mt_mix(mt[i], mt[i + 1], mt[i + kMtM]);
    __syncthreads();
    if (i < kMtN - kMtM) mt[i] = v;
    __syncthreads();
  }
  for (int b = kMtN - kMtM; b < 2 * (kMtN - kMtM); b += 64) {  // [227, 454)
    const int i = b + lane;
    uint32_t v = 0;
    if (i < 2 * (kMtN - kMtM))
      v = mt_mix(mt[i], mt[i + 1], mt[i + kMtM - kMtN]);
    __syncthreads();
    if (i < 2 * (kMtN - kMtM)) mt[i] = v;
    __syncthreads();
  }
  for (int b = 2 * (kMtN - kMtM); b < kMtN - 1; b += 64) {  // [454, 623)
    const int i = b + lane;
    uint32_t v = 0;
    if (i < kMtN - 1) v = mt_mix(mt[i], mt[i + 1], mt[i + kMtM - kMtN]);
    __syncthreads();
    if (i < kMtN - 1) mt[i] = v;
    __syncthreads();
  }
  uint32_t last = 0;
  if (lane == 0) last = mt_mix(mt[kMtN - 1], mt[0], mt[kMtM - 1]);
  __syncthreads();
  if (lane == 0) mt[kMtN - 1] = last;
  __syncthreads();
}

__device__ inline uint32_t mt_temper(uint32_t y) {
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

__device__ inline uint32_t mt_next(uint32_t* mt, int& mti, int lane) {
  if (mti >= kMtN) {
    mt_twist(mt, lane);
    mti = 0;
  }
  return mt_temper(mt[mti++]);
}

// The next four words (two random() draws).  The common case reads them
// straight from the block; near the block's end a loop with the one inlined
// regeneration takes them one by one.
__device__ inline void mt_next4(uint32_t* mt, int& mti, int lane,
                                uint32_t w[4]) {
  if (mti <= kMtN - 4) {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = mt_temper(mt[mti + k]);
    mti += 4;
    return;
  }
#pragma unroll 1
  for (int k = 0; k < 4; ++k) {
    if (mti >= kMtN) {
      mt_twist(mt, lane);
      mti = 0;
    }
    w[k] = mt_temper(mt[mti++]);
  }
}

__device__ inline double mt_uniform(uint32_t a, uint32_t b) {
  return (double(a >> 5) * 67108864.0 + double(b >> 6)) *
         (1.0 / 9007199254740992.0);
}

}  // namespace pycd

// Reduced-precision reciprocal contraction of the Ewald table on MFMA.
//
// A_recip = 2/eps * Y Y^T with the structure factors Y (N x 2K) is computed
// on the matrix cores in one of two modes:
//   F32     Y in FP32, v_mfma_f32_16x16x4f32, FP32 accumulation;
//   BF16X3  Y = hi + lo with hi = bf16(Y), lo = bf16(Y - hi) (both from the
//           FP64 structure factor), Y Y^T ~ hi hi^T + hi lo^T + lo hi^T on
//           v_mfma_f32_16x16x32_bf16, FP32 accumulation.
// The real-space term (minimum image, erfc), the self term and the final
// FP64 store are the FP64 epilogue shared with the FP64 kernel, so only the
// reciprocal sum carries the reduced precision.
//
// Tiling: 256 threads = 4 wavefronts per 128 x 128 output tile; each
// wavefront owns 64 x 64 = 4 x 4 MFMA blocks of 16 x 16.  K is staged
// through LDS in slices of 32 (row-major, padded rows); the global loads of
// the next slice are in flight while the current one is multiplied.  The
// BF16X3 mode also has two-stage kernels (one barrier per slice, loads two
// or three slices ahead) on 128 x 128 tiles and on 256 x 256 tiles of 8
// wavefronts; the latter is the default from 4096 sites on.
#include "ewald_common.hpp"

#include <hip/hip_bf16.h>

namespace pycd {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kMTile = 128;
constexpr int kMSlice = 32;

// 16-byte load through a global (address-space-1) pointer.  A generic
// pointer would compile to a flat load, which also counts against the LDS
// wait counter and would make every LDS read wait for the prefetch.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ inline uint4 load16_global(const void* p) {
  const u32x4 v = *(const __attribute__((address_space(1))) u32x4*)(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ inline float4 load16f_global(const void* p) {
  const uint4 v = load16_global(p);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y),
                     __uint_as_float(v.z), __uint_as_float(v.w));
}

// Structure factors in reduced precision from the FP64 values.
__global__ void __launch_bounds__(256)
structure_factor_lowp(const double* __restrict__ pos, int n,
                      const double* __restrict__ kv,
                      const double* __restrict__ sw, int K, int Kp, int mode,
                      float* __restrict__ y32, __bf16* __restrict__ yhi,
                      __bf16* __restrict__ ylo) {
  const int64_t total = int64_t(n) * K;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(idx / K);
    const int m = int(idx - int64_t(i) * K);
    const double ph = kv[3 * m] * pos[3 * i] + kv[3 * m + 1] * pos[3 * i + 1] +
                      kv[3 * m + 2] * pos[3 * i + 2];
    double s, c;
    sincos(ph, &s, &c);
    const double vc = sw[m] * c, vs = sw[m] * s;
    const int64_t oc = int64_t(i) * Kp + m, os = int64_t(i) * Kp + K + m;
    if (mode == PYCD_PREC_F32) {
      y32[oc] = float(vc);
      y32[os] = float(vs);
    } else {
      const __bf16 hc = __bf16(float(vc)), hs = __bf16(float(vs));
      yhi[oc] = hc;
      yhi[os] = hs;
      ylo[oc] = __bf16(float(vc - double(float(hc))));
      ylo[os] = __bf16(float(vs - double(float(hs))));
    }
  }
}

__global__ void __launch_bounds__(256)
table_mfma_f32(const float* __restrict__ Y, int n, int Kp,
               const double* __restrict__ pos, MicGeometry g, EwaldScalars es,
               double* __restrict__ A, unsigned long long* __restrict__ pairs) {
  __shared__ float As[kMTile][kMSlice + 1];
  __shared__ float Bs[kMTile][kMSlice + 1];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int bi, bj;
  tile_coords_upper(n, kMTile, bi, bj);
  const int i0 = bi * kMTile, j0 = bj * kMTile;
  const bool mirror = bi != bj;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  const int lrow = t >> 1, lk = (t & 1) * 16;
  const bool a_ok = i0 + lrow < n, b_ok = j0 + lrow < n;
  const float* ya = Y + int64_t(i0 + lrow) * Kp + lk;
  const float* yb = Y + int64_t(j0 + lrow) * Kp + lk;
  // software pipeline: slice k0 + kMSlice is loaded into registers while the
  // matrix cores work on slice k0 from LDS
  float4 va[4], vb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      va[q] = a_ok ? load16f_global(ya + k0 + 4 * q)
                   : make_float4(0.f, 0.f, 0.f, 0.f);
      vb[q] = b_ok ? load16f_global(yb + k0 + 4 * q)
                   : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[lrow][lk + 4 * q + 0] = va[q].x;
      As[lrow][lk + 4 * q + 1] = va[q].y;
      As[lrow][lk + 4 * q + 2] = va[q].z;
      As[lrow][lk + 4 * q + 3] = va[q].w;
      Bs[lrow][lk + 4 * q + 0] = vb[q].x;
      Bs[lrow][lk + 4 * q + 1] = vb[q].y;
      Bs[lrow][lk + 4 * q + 2] = vb[q].z;
      Bs[lrow][lk + 4 * q + 3] = vb[q].w;
    }
  };
  gload(0);
  sstore();
  __syncthreads();
  for (int k0 = 0; k0 < Kp; k0 += kMSlice) {
    const bool more = k0 + kMSlice < Kp;
    if (more) gload(k0 + kMSlice);
#pragma unroll
    for (int kk = 0; kk < kMSlice; kk += 4) {
      float a[4], b[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        a[x] = As[wr + 16 * x + (lane & 15)][kk + (lane >> 4)];
        b[x] = Bs[wc + 16 * x + (lane & 15)][kk + (lane >> 4)];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[x], b[y],
                                                           acc[x][y], 0, 0, 0);
    }
    __syncthreads();  // every wavefront is done with this slice
    if (more) sstore();
    __syncthreads();
  }
  unsigned cnt = 0;
  // entry (i, j) and, off the diagonal block, its mirror (j, i): the same
  // number, formed once (store_pair)
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + 16 * x + (lane >> 4) * 4 + r;
        const int j = j0 + wc + 16 * y + (lane & 15);
        if (i < n && j < n)
          store_pair(A, n, i, j, mirror, pos, g, es, double(acc[x][y][r]),
                     cnt);
      }
  count_pairs(cnt, pairs);
}

// K is staged in slices of kBSlice = 32 or 64: a 64-deep slice makes two
// 32-deep MFMA steps per pair of barriers, and its global loads (in
// registers while the previous slice is multiplied) have twice the matrix
// work to arrive in.  Padded rows of kBSlice + 8 bf16 put the 8 rows one
// 16-byte LDS read serves on distinct banks.  The sums over k run in the
// same order for both, so they give the same table bit for bit.
template <int kBSlice>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(2, 2)))
table_mfma_bf16x3(const __bf16* __restrict__ Yhi,
                  const __bf16* __restrict__ Ylo, int n, int Kp,
                  const double* __restrict__ pos, MicGeometry g,
                  EwaldScalars es, double* __restrict__ A,
                  unsigned long long* __restrict__ pairs) {
  constexpr int kRow = kBSlice + 8;  // padded row (bf16 elements)
  __shared__ __attribute__((aligned(16))) __bf16 Ah[kMTile * kRow];
  __shared__ __attribute__((aligned(16))) __bf16 Al[kMTile * kRow];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[kMTile * kRow];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[kMTile * kRow];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int bi, bj;
  tile_coords_upper(n, kMTile, bi, bj);
  const int i0 = bi * kMTile, j0 = bj * kMTile;
  const bool mirror = bi != bj;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  // loader: 128 rows x kBSlice bf16 = 256 threads x kBSlice / 2 bf16
  // (16-byte chunks) per array
  constexpr int kQ = kBSlice / 16;  // chunks per thread and array
  const int lrow = t >> 1, lk = (t & 1) * (kBSlice / 2);
  const bool a_ok = i0 + lrow < n, b_ok = j0 + lrow < n;
  const int64_t aoff = int64_t(i0 + lrow) * Kp + lk;
  const int64_t boff = int64_t(j0 + lrow) * Kp + lk;
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  uint4 ah[kQ], al[kQ], bh[kQ], bl[kQ];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      ah[q] = a_ok ? load16_global(Yhi + aoff + k0 + 8 * q) : zero;
      al[q] = a_ok ? load16_global(Ylo + aoff + k0 + 8 * q) : zero;
      bh[q] = b_ok ? load16_global(Yhi + boff + k0 + 8 * q) : zero;
      bl[q] = b_ok ? load16_global(Ylo + boff + k0 + 8 * q) : zero;
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      *reinterpret_cast<uint4*>(Ah + lrow * kRow + lk + 8 * q) = ah[q];
      *reinterpret_cast<uint4*>(Al + lrow * kRow + lk + 8 * q) = al[q];
      *reinterpret_cast<uint4*>(Bh + lrow * kRow + lk + 8 * q) = bh[q];
      *reinterpret_cast<uint4*>(Bl + lrow * kRow + lk + 8 * q) = bl[q];
    }
  };
  gload(0);
  sstore();
  __syncthreads();
  for (int k0 = 0; k0 < Kp; k0 += kBSlice) {
    const bool more = k0 + kBSlice < Kp;
    if (more) gload(k0 + kBSlice);
#pragma unroll
    for (int h = 0; h < kBSlice / 32; ++h) {
      // this wavefront's column blocks once, then one row block at a time
      const int kof = (lane >> 4) * 8 + 32 * h;
      v8bf b_h[4], b_l[4];
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const int rb = (wc + 16 * y + (lane & 15)) * kRow + kof;
        b_h[y] = *reinterpret_cast<const v8bf*>(Bh + rb);
        b_l[y] = *reinterpret_cast<const v8bf*>(Bl + rb);
      }
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int ra = (wr + 16 * x + (lane & 15)) * kRow + kof;
        const v8bf a_h = *reinterpret_cast<const v8bf*>(Ah + ra);
        const v8bf a_l = *reinterpret_cast<const v8bf*>(Al + ra);
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              a_l, b_h[y], acc[x][y], 0, 0, 0);
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              a_h, b_l[y], acc[x][y], 0, 0, 0);
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              a_h, b_h[y], acc[x][y], 0, 0, 0);
        }
      }
    }
    __syncthreads();  // every wavefront is done with this slice
    if (more) sstore();
    __syncthreads();
  }
  unsigned cnt = 0;
  // entry (i, j) and, off the diagonal block, its mirror (j, i): the same
  // number, formed once (store_pair)
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + 16 * x + (lane >> 4) * 4 + r;
        const int j = j0 + wc + 16 * y + (lane & 15);
        if (i < n && j < n)
          store_pair(A, n, i, j, mirror, pos, g, es, double(acc[x][y][r]),
                     cnt);
      }
  count_pairs(cnt, pairs);
}

// K slices of 32 in two LDS stages and three register sets: while slice t
// is multiplied out of one stage, slice t + 1 (in registers since two steps
// before) goes to the other stage and slice t + 3 is loaded into the
// registers slice t left, so a global load has three steps to arrive.  One barrier per slice, without a
// memory fence: it waits for this wavefront's LDS operations only, and the
// loads in flight stay in flight.  Two workgroups fill a CU's 160 KB of LDS.
// The sums over k run in the same order as the one-stage kernels, so the
// table is the same bit for bit.  Y holds whole 128-row tiles (rows past n
// are zero) and Kp is a multiple of 6 kB (zero columns), so every load and
// LDS store is made unconditionally: a branch around them would make the
// compiler wait for all loads in flight.
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(2, 2)))
table_mfma_bf16x3_2stage(const __bf16* __restrict__ Yhi,
                         const __bf16* __restrict__ Ylo, int n, int Kp,
                         const double* __restrict__ pos, MicGeometry g,
                         EwaldScalars es, double* __restrict__ A,
                         unsigned long long* __restrict__ pairs) {
  constexpr int kB = 32;
  constexpr int kRow = kB + 8;  // padded row (bf16 elements)
  constexpr int kArr = kMTile * kRow;
  // stage s, array a (hi A, lo A, hi B, lo B) at S + (4 s + a) kArr
  __shared__ __attribute__((aligned(16))) __bf16 S[2 * 4 * kArr];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int bi, bj;
  tile_coords_upper(n, kMTile, bi, bj);
  const int i0 = bi * kMTile, j0 = bj * kMTile;
  const bool mirror = bi != bj;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  // loader: 128 rows x 32 bf16 per array = 256 threads x one 16-byte chunk
  // of each of hi/lo at two k offsets
  constexpr int kQ = kB / 16;
  const int lrow = t >> 1, lk = (t & 1) * (kB / 2);
  const int64_t aoff = int64_t(i0 + lrow) * Kp + lk;
  const int64_t boff = int64_t(j0 + lrow) * Kp + lk;
  struct Regs { uint4 v[4][kQ]; };
  auto gload = [&](Regs& r, int k0) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      r.v[0][q] = load16_global(Yhi + aoff + k0 + 8 * q);
      r.v[1][q] = load16_global(Ylo + aoff + k0 + 8 * q);
      r.v[2][q] = load16_global(Yhi + boff + k0 + 8 * q);
      r.v[3][q] = load16_global(Ylo + boff + k0 + 8 * q);
    }
  };
  auto sstore = [&](int stage, const Regs& r) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int q = 0; q < kQ; ++q)
        *reinterpret_cast<uint4*>(S + (4 * stage + a) * kArr + lrow * kRow +
                                  lk + 8 * q) = r.v[a][q];
  };
  // all of this wavefront's LDS operations done, then the barrier
  auto lds_barrier = [] {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };
  auto multiply = [&](int stage) {
    const __bf16* Ah = S + (4 * stage + 0) * kArr;
    const __bf16* Al = S + (4 * stage + 1) * kArr;
    const __bf16* Bh = S + (4 * stage + 2) * kArr;
    const __bf16* Bl = S + (4 * stage + 3) * kArr;
    const int kof = (lane >> 4) * 8;
    v8bf b_h[4], b_l[4];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int rb = (wc + 16 * y + (lane & 15)) * kRow + kof;
      b_h[y] = *reinterpret_cast<const v8bf*>(Bh + rb);
      b_l[y] = *reinterpret_cast<const v8bf*>(Bl + rb);
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int ra = (wr + 16 * x + (lane & 15)) * kRow + kof;
      const v8bf a_h = *reinterpret_cast<const v8bf*>(Ah + ra);
      const v8bf a_l = *reinterpret_cast<const v8bf*>(Al + ra);
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_l, b_h[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_h, b_l[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_h, b_h[y], acc[x][y], 0, 0, 0);
      }
    }
  };
  // step for slice k0 (in stage `stage`): `next` holds slice k0 + kB,
  // `spare` takes slice k0 + 3 kB; at the end of K they load and store the
  // last slice again (into a stage no step reads)
  // (scheduling barriers keep the loads at the start of the step and the
  // LDS stores at its end, which the scheduler would otherwise swap)
  auto step = [&](int stage, int k0, Regs& next, Regs& spare) {
    gload(spare, min(k0 + 3 * kB, Kp - kB));
    __builtin_amdgcn_sched_barrier(0);
    multiply(stage);
    __builtin_amdgcn_sched_barrier(0);
    sstore(stage ^ 1, next);
    lds_barrier();
  };
  Regs r0, r1, r2;
  gload(r0, 0);
  sstore(0, r0);
  gload(r1, kB);
  gload(r2, 2 * kB);
  lds_barrier();
  for (int k0 = 0; k0 < Kp; k0 += 6 * kB) {
    step(0, k0, r1, r0);
    step(1, k0 + kB, r2, r1);
    step(0, k0 + 2 * kB, r0, r2);
    step(1, k0 + 3 * kB, r1, r0);
    step(0, k0 + 4 * kB, r2, r1);
    step(1, k0 + 5 * kB, r0, r2);
  }
  unsigned cnt = 0;
  // entry (i, j) and, off the diagonal block, its mirror (j, i): the same
  // number, formed once (store_pair)
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + 16 * x + (lane >> 4) * 4 + r;
        const int j = j0 + wc + 16 * y + (lane & 15);
        if (i < n && j < n)
          store_pair(A, n, i, j, mirror, pos, g, es, double(acc[x][y][r]),
                     cnt);
      }
  count_pairs(cnt, pairs);
}

// 256 x 256 output tiles on 8 wavefronts (512 threads), each wavefront
// 128 x 64 = 8 x 4 MFMA blocks: a step loads half the bytes per product of
// the 128 x 128 tiles, and a wavefront reads a quarter less LDS per
// product.  K slices of 32 in two LDS stages (160 KB, one workgroup per CU):
// a step loads the next slice into registers as it starts and stores it to
// the other stage as it ends, with the step's 96 MFMAs per wavefront in
// between (one register set: the accumulators take half the registers).  For the table to equal
// the 128 x 128 kernels' bit for bit, every entry is formed with the same
// operands in the same roles: in a diagonal 256-tile the lower-left 128 x 128
// block is not stored, its upper mirror writes it (as a 128 x 128 tile off
// the diagonal does), and the two diagonal 128-blocks store both (i, j) and
// (j, i) as formed.  Y holds whole 256-row tiles and Kp is a multiple of
// 2 kB, so loads and LDS stores are unconditional.  The real-space part of
// each entry it forms is in A already (table_real_part), so the epilogue
// adds the reciprocal and self terms to it: the same operations, in the
// same order, as store_pair.
constexpr int kMTile2 = 256;

// Two 128 x 128 tiles in one 512-thread workgroup of the 256 x 256 kernel,
// one per half (4 wavefronts of 64 x 64, as in the 128 x 128 kernels), each
// half in its own 80 KB of the LDS, the two in step at the shared
// barriers.  `sub` lists the tiles (tile row, tile column); workgroup k
// takes entries 2k and 2k + 1 (a missing second one repeats the first and
// stores nothing).
__device__ __forceinline__ void bf16x3_tile_pair(
    __bf16* S, const __bf16* __restrict__ Yhi, const __bf16* __restrict__ Ylo,
    int n, int Kp, const EwaldScalars& es, double* __restrict__ A,
    const int2* __restrict__ sub, int n_sub, int k) {
  constexpr int kB = 32;
  constexpr int kRow = kB + 8;
  constexpr int kArr = kMTile * kRow;
  const int t = threadIdx.x, th = t & 255, lane = t & 63;
  const int half = __builtin_amdgcn_readfirstlane(t >> 8);
  const int wave = __builtin_amdgcn_readfirstlane((t >> 6) & 3);
  int idx = 2 * k + half;
  const bool dummy = idx >= n_sub;
  if (dummy) idx = 2 * k;
  const int2 tl = sub[idx];
  const int i0 = tl.x * kMTile, j0 = tl.y * kMTile;
  const bool mirror = tl.x != tl.y;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  __bf16* Sh = S + half * (2 * 4 * kArr);
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  constexpr int kQ = kB / 16;
  const int lrow = th >> 1, lk = (th & 1) * (kB / 2);
  const int64_t aoff = int64_t(i0 + lrow) * Kp + lk;
  const int64_t boff = int64_t(j0 + lrow) * Kp + lk;
  struct Regs { uint4 v[4][kQ]; };
  auto gload = [&](Regs& r, int k0) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      r.v[0][q] = load16_global(Yhi + aoff + k0 + 8 * q);
      r.v[1][q] = load16_global(Ylo + aoff + k0 + 8 * q);
      r.v[2][q] = load16_global(Yhi + boff + k0 + 8 * q);
      r.v[3][q] = load16_global(Ylo + boff + k0 + 8 * q);
    }
  };
  auto sstore = [&](int stage, const Regs& r) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int q = 0; q < kQ; ++q)
        *reinterpret_cast<uint4*>(Sh + (4 * stage + a) * kArr + lrow * kRow +
                                  lk + 8 * q) = r.v[a][q];
  };
  auto lds_barrier = [] {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };
  auto multiply = [&](int stage) {
    const __bf16* Ah = Sh + (4 * stage + 0) * kArr;
    const __bf16* Al = Sh + (4 * stage + 1) * kArr;
    const __bf16* Bh = Sh + (4 * stage + 2) * kArr;
    const __bf16* Bl = Sh + (4 * stage + 3) * kArr;
    const int kof = (lane >> 4) * 8;
    v8bf b_h[4], b_l[4];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int rb = (wc + 16 * y + (lane & 15)) * kRow + kof;
      b_h[y] = *reinterpret_cast<const v8bf*>(Bh + rb);
      b_l[y] = *reinterpret_cast<const v8bf*>(Bl + rb);
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int ra = (wr + 16 * x + (lane & 15)) * kRow + kof;
      const v8bf a_h = *reinterpret_cast<const v8bf*>(Ah + ra);
      const v8bf a_l = *reinterpret_cast<const v8bf*>(Al + ra);
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_l, b_h[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_h, b_l[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_h, b_h[y], acc[x][y], 0, 0, 0);
      }
    }
  };
  Regs r;
  auto step = [&](int stage, int k0) {
    gload(r, min(k0 + kB, Kp - kB));
    __builtin_amdgcn_sched_barrier(0);
    multiply(stage);
    __builtin_amdgcn_sched_barrier(0);
    sstore(stage ^ 1, r);
    lds_barrier();
  };
  gload(r, 0);
  sstore(0, r);
  lds_barrier();
  for (int k0 = 0; k0 < Kp; k0 += 2 * kB) {
    step(0, k0);
    step(1, k0 + kB);
  }
  if (dummy) return;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const int i = i0 + wr + 16 * x + (lane >> 4) * 4 + r4;
        const int j = j0 + wc + 16 * y + (lane & 15);
        if (i < n && j < n) {
          const double v = finish_entry(i, j, es, A[int64_t(i) * n + j],
                                        double(acc[x][y][r4]));
          A[int64_t(i) * n + j] = v;
          if (mirror) A[int64_t(j) * n + i] = v;
        }
      }
}

// Real-space part of every entry that a pass over 128 x 128 tiles forms as
// (i, j) (tile row <= tile column), into A ahead of the contraction, with
// the pairs within r_cut counted as store_pair counts them (a tile off the
// diagonal also for its mirror).  One 256-thread workgroup per tile, full
// occupancy: the minimum image and erfc run while no matrix work waits.
__global__ void __launch_bounds__(256)
table_real_part(const double* __restrict__ pos, int n, MicGeometry g,
                EwaldScalars es, double* __restrict__ A,
                unsigned long long* __restrict__ pairs) {
  int bi, bj;
  tile_coords_upper(n, kMTile, bi, bj);
  const int i0 = bi * kMTile, j0 = bj * kMTile;
  const bool mirror = bi != bj;
  // a lane keeps its column j; a wavefront walks rows i the same for all
  // of its lanes (every other row, the two halves of the workgroup
  // interleaved), so the row positions come through scalar loads
  const int j = j0 + (threadIdx.x & (kMTile - 1));
  const int ih = __builtin_amdgcn_readfirstlane(threadIdx.x / kMTile);
  const bool j_ok = j < n;
  const double pjx = j_ok ? pos[3 * j] : 0.0;
  const double pjy = j_ok ? pos[3 * j + 1] : 0.0;
  const double pjz = j_ok ? pos[3 * j + 2] : 0.0;
  const int i_end = min(kMTile, n - i0);
  unsigned cnt = 0;
  for (int k = ih; k < i_end; k += 2) {
    const int i = i0 + k;
    if (j_ok) {
      const unsigned before = cnt;
      A[int64_t(i) * n + j] =
          real_part_d(i, j, pjx - pos[3 * i], pjy - pos[3 * i + 1],
                      pjz - pos[3 * i + 2], g, es, cnt);
      if (mirror) cnt += cnt - before;
    }
  }
  count_pairs(cnt, pairs);
}

__global__ void __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(2, 2)))
table_mfma_bf16x3_256(const __bf16* __restrict__ Yhi,
                      const __bf16* __restrict__ Ylo, int n, int Kp,
                      EwaldScalars es, double* __restrict__ A,
                      int main_tiles, const int2* __restrict__ sub,
                      int n_sub) {
  constexpr int kB = 32;
  constexpr int kRow = kB + 8;  // padded row (bf16 elements)
  constexpr int kArr = kMTile2 * kRow;
  // stage s, array a (hi A, lo A, hi B, lo B) at S + (4 s + a) kArr
  __shared__ __attribute__((aligned(16))) __bf16 S[2 * 4 * kArr];
  // the tiles past the last whole round of 256 x 256 tiles, as pairs of
  // 128 x 128 tiles (the last round takes about half as long)
  if (int(blockIdx.x) >= main_tiles) {
    bf16x3_tile_pair(S, Yhi, Ylo, n, Kp, es, A, sub, n_sub,
                     int(blockIdx.x) - main_tiles);
    return;
  }
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  int bi, bj;
  tile_coords_upper(n, kMTile2, bi, bj);
  const int i0 = bi * kMTile2, j0 = bj * kMTile2;
  const int wr = (wave >> 2) * 128, wc = (wave & 3) * 64;
  // in a diagonal tile: the upper-right 128-block is stored with its mirror,
  // the lower-left one not at all
  const bool diag = bi == bj;
  const bool mirror = !diag || (wr == 0 && wc >= 128);
  const bool skip = diag && wr == 128 && wc < 128;
  v4f acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  // loader: 256 rows x 32 bf16 per array = 512 threads x two 16-byte chunks
  constexpr int kQ = kB / 16;
  const int lrow = t >> 1, lk = (t & 1) * (kB / 2);
  const int64_t aoff = int64_t(i0 + lrow) * Kp + lk;
  const int64_t boff = int64_t(j0 + lrow) * Kp + lk;
  struct Regs { uint4 v[4][kQ]; };
  auto gload = [&](Regs& r, int k0) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      r.v[0][q] = load16_global(Yhi + aoff + k0 + 8 * q);
      r.v[1][q] = load16_global(Ylo + aoff + k0 + 8 * q);
      r.v[2][q] = load16_global(Yhi + boff + k0 + 8 * q);
      r.v[3][q] = load16_global(Ylo + boff + k0 + 8 * q);
    }
  };
  auto sstore = [&](int stage, const Regs& r) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int q = 0; q < kQ; ++q)
        *reinterpret_cast<uint4*>(S + (4 * stage + a) * kArr + lrow * kRow +
                                  lk + 8 * q) = r.v[a][q];
  };
  auto lds_barrier = [] {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };
  auto multiply = [&](int stage) {
    const __bf16* Ah = S + (4 * stage + 0) * kArr;
    const __bf16* Al = S + (4 * stage + 1) * kArr;
    const __bf16* Bh = S + (4 * stage + 2) * kArr;
    const __bf16* Bl = S + (4 * stage + 3) * kArr;
    const int kof = (lane >> 4) * 8;
    v8bf b_h[4], b_l[4];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int rb = (wc + 16 * y + (lane & 15)) * kRow + kof;
      b_h[y] = *reinterpret_cast<const v8bf*>(Bh + rb);
      b_l[y] = *reinterpret_cast<const v8bf*>(Bl + rb);
    }
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      const int ra = (wr + 16 * x + (lane & 15)) * kRow + kof;
      const v8bf a_h = *reinterpret_cast<const v8bf*>(Ah + ra);
      const v8bf a_l = *reinterpret_cast<const v8bf*>(Al + ra);
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_l, b_h[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_h, b_l[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            a_h, b_h[y], acc[x][y], 0, 0, 0);
      }
    }
  };
  // step for slice k0 (in stage `stage`): slice k0 + kB is loaded and
  // goes to the other stage; the last step loads and stores the last slice
  // again (into a stage no step reads); scheduling barriers keep the loads
  // at the start of the step and the LDS stores at its end
  Regs r;
  auto step = [&](int stage, int k0) {
    gload(r, min(k0 + kB, Kp - kB));
    __builtin_amdgcn_sched_barrier(0);
    multiply(stage);
    __builtin_amdgcn_sched_barrier(0);
    sstore(stage ^ 1, r);
    lds_barrier();
  };
  gload(r, 0);
  sstore(0, r);
  lds_barrier();
  for (int k0 = 0; k0 < Kp; k0 += 2 * kB) {
    step(0, k0);
    step(1, k0 + kB);
  }
  if (skip) return;
#pragma unroll
  for (int x = 0; x < 8; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wr + 16 * x + (lane >> 4) * 4 + r;
        const int j = j0 + wc + 16 * y + (lane & 15);
        if (i < n && j < n) {
          const double v = finish_entry(i, j, es, A[int64_t(i) * n + j],
                                        double(acc[x][y][r]));
          A[int64_t(i) * n + j] = v;
          if (mirror) A[int64_t(j) * n + i] = v;
        }
      }
}

int launch_table_mfma(pycd_ctx* ctx, int precision, const double* pos, int n,
                      const double* kv, const double* sw, int K,
                      const MicGeometry& geo, const EwaldScalars& es,
                      double* out, unsigned long long* pairs,
                      hipEvent_t after_sf, hipEvent_t after_real,
                      bool& real_apart) {
  hipStream_t s = ctx->stream;
  // K padded to whole slices of the mode's kernel (zero columns add nothing)
  // BF16x3 default: 256 x 256 tiles from 4096 sites on, K slices of 64
  // below (a few large tiles would leave most CUs idle)
  int bvar = ctx->table_variant[PYCD_PREC_BF16X3];
  if (bvar == 0) bvar = n >= 4096 ? 4 : 2;
  const int bslice = bvar == 1 || bvar >= 3 ? 32 : 64;
  // (the two-stage kernels take K in groups of six and of two 32-slices)
  const int slice = precision == PYCD_PREC_F32 ? kMSlice
                    : bvar == 3              ? 6 * bslice
                    : bvar == 4              ? 2 * bslice
                                             : bslice;
  const int Kp = std::max(slice, ((2 * K + slice - 1) / slice) * slice);
  // rows in whole tiles: the rows past n stay zero
  const int tile = precision == PYCD_PREC_BF16X3 && bvar == 4 ? kMTile2
                                                              : kMTile;
  const int nb = (n + tile - 1) / tile;
  const size_t elems = size_t(nb) * tile * Kp;
  DevBuf<float> y32;
  DevBuf<__bf16> yhi, ylo;
  if (precision == PYCD_PREC_F32) {
    PYCD_HIP(ctx, y32.alloc(elems));
    PYCD_HIP(ctx, hipMemsetAsync(y32.ptr, 0, elems * sizeof(float), s));
  } else {
    PYCD_HIP(ctx, yhi.alloc(elems));
    PYCD_HIP(ctx, ylo.alloc(elems));
    PYCD_HIP(ctx, hipMemsetAsync(yhi.ptr, 0, elems * sizeof(__bf16), s));
    PYCD_HIP(ctx, hipMemsetAsync(ylo.ptr, 0, elems * sizeof(__bf16), s));
  }
  if (K > 0) {
    const int64_t total = int64_t(n) * K;
    const int blocks = int(std::min<int64_t>((total + 255) / 256, 65536));
    structure_factor_lowp<<<blocks, 256, 0, s>>>(pos, n, kv, sw, K, Kp,
                                                 precision, y32.ptr, yhi.ptr,
                                                 ylo.ptr);
    PYCD_HIP(ctx, hipGetLastError());
  }
  PYCD_HIP(ctx, hipEventRecord(after_sf, s));
  // the real-space term has a kernel of its own (timed) only ahead of the
  // 256 x 256-tile contraction; elsewhere it is in the epilogue
  real_apart = precision == PYCD_PREC_BF16X3 && bvar == 4;
  if (!real_apart) PYCD_HIP(ctx, hipEventRecord(after_real, s));
  const int tiles = nb * (nb + 1) / 2;
  if (precision == PYCD_PREC_F32)
    table_mfma_f32<<<tiles, 256, 0, s>>>(y32.ptr, n, Kp, pos, geo, es, out,
                                         pairs);
  else if (bvar == 3)
    table_mfma_bf16x3_2stage<<<tiles, 256, 0, s>>>(yhi.ptr, ylo.ptr, n, Kp,
                                                   pos, geo, es, out, pairs);
  else if (bvar == 4) {
    const int nb1 = (n + kMTile - 1) / kMTile;
    table_real_part<<<nb1 * (nb1 + 1) / 2, 256, 0, s>>>(pos, n, geo, es, out,
                                                        pairs);
    PYCD_HIP(ctx, hipGetLastError());
    PYCD_HIP(ctx, hipEventRecord(after_real, s));
    // whole rounds of 256 x 256 tiles (one workgroup per CU), the rest as
    // pairs of 128 x 128 tiles; fewer tiles than one round all go whole
    int dev_cus = 0;
    PYCD_HIP(ctx, hipDeviceGetAttribute(&dev_cus,
                                        hipDeviceAttributeMultiprocessorCount,
                                        ctx->device));
    const int round = std::max(dev_cus, 1);
    const int main_tiles = tiles >= round ? (tiles / round) * round : tiles;
    std::vector<int2> sub;
    for (int w = main_tiles; w < tiles; ++w) {
      int bi = 0, bj = 0;
      tile_coords_upper_at(n, kMTile2, w, bi, bj);
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          const int si = 2 * bi + a, sj = 2 * bj + b;
          if (si > sj || si * kMTile >= n || sj * kMTile >= n) continue;
          sub.push_back(make_int2(si, sj));
        }
    }
    DevBuf<int2> d_sub;
    if (!sub.empty()) {
      PYCD_HIP(ctx, d_sub.alloc(sub.size()));
      PYCD_HIP(ctx, hipMemcpyAsync(d_sub.ptr, sub.data(),
                                   sizeof(int2) * sub.size(),
                                   hipMemcpyHostToDevice, s));
    }
    const int grid = main_tiles + int((sub.size() + 1) / 2);
    table_mfma_bf16x3_256<<<grid, 512, 0, s>>>(yhi.ptr, ylo.ptr, n, Kp, es,
                                               out, main_tiles, d_sub.ptr,
                                               int(sub.size()));
    PYCD_HIP(ctx, hipGetLastError());
    PYCD_HIP(ctx, hipStreamSynchronize(s));
  }
  else if (bslice == 32)
    table_mfma_bf16x3<32><<<tiles, 256, 0, s>>>(yhi.ptr, ylo.ptr, n, Kp, pos,
                                                geo, es, out, pairs);
  else
    table_mfma_bf16x3<64><<<tiles, 256, 0, s>>>(yhi.ptr, ylo.ptr, n, Kp, pos,
                                                geo, es, out, pairs);
  PYCD_HIP(ctx, hipGetLastError());
  PYCD_HIP(ctx, hipStreamSynchronize(s));
  return PYCD_OK;
}

}  // namespace pycd

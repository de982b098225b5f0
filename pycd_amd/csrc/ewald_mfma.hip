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
// the next slice are in flight while the current one is multiplied.
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

int launch_table_mfma(pycd_ctx* ctx, int precision, const double* pos, int n,
                      const double* kv, const double* sw, int K,
                      const MicGeometry& geo, const EwaldScalars& es,
                      double* out, unsigned long long* pairs,
                      hipEvent_t after_sf) {
  hipStream_t s = ctx->stream;
  // K padded to whole slices of the mode's kernel (zero columns add nothing)
  const int bvar = ctx->table_variant[PYCD_PREC_BF16X3];
  const int bslice = bvar == 1 ? 32 : 64;
  const int slice = precision == PYCD_PREC_F32 ? kMSlice : bslice;
  const int Kp = std::max(slice, ((2 * K + slice - 1) / slice) * slice);
  const size_t elems = size_t(n) * Kp;
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
  const int nb = (n + kMTile - 1) / kMTile;
  const int tiles = nb * (nb + 1) / 2;
  if (precision == PYCD_PREC_F32)
    table_mfma_f32<<<tiles, 256, 0, s>>>(y32.ptr, n, Kp, pos, geo, es, out,
                                         pairs);
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

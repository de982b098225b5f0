// Ewald pair table of PyCD's System.get_precomputed_array on gfx950.
//
//   A[i,j] = erfc(sqrt(alpha) r_ij) / (2 r_ij) / eps      (r_ij < r_cut)
//          + 2/eps * sum_{k in half space, |k|^2 < k_cut^2}
//                    (2 pi / V) exp(-k^2 / (4 alpha)) / k^2 cos(k . r_ij)
//          - sqrt(alpha / pi) / eps * delta_ij
// with the real-space diagonal equal to 0.5/eps (PyCD/core.py:799-814,
// 853-878, 1594-1661).
//
// The reciprocal sum is recast as a symmetric rank-2K product of per-site
// structure factors Y[i, m] = sqrt(w_m) cos(k_m . r_i),
// Y[i, K + m] = sqrt(w_m) sin(k_m . r_i): A_recip = 2/eps * Y Y^T, because
// cos(k.(r_j - r_i)) = cos(k.r_i) cos(k.r_j) + sin(k.r_i) sin(k.r_j).  The
// contraction kernel fuses the real-space term (minimum image + erfc), the
// self term and the final store, so the N x N table is written once.
#include "common.hpp"
#include "ewald_common.hpp"

#include <cmath>

namespace pycd {

// --------------------------------------------------------- structure factors
__global__ void __launch_bounds__(256)
structure_factor_f64(const double* __restrict__ pos, int n,
                     const double* __restrict__ kv, const double* __restrict__ sw,
                     int K, int Kp, double* __restrict__ Y) {
  const int64_t total = int64_t(n) * K;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(idx / K);
    const int m = int(idx - int64_t(i) * K);
    const double ph = kv[3 * m] * pos[3 * i] + kv[3 * m + 1] * pos[3 * i + 1] +
                      kv[3 * m + 2] * pos[3 * i + 2];
    double s, c;
    sincos(ph, &s, &c);
    Y[int64_t(i) * Kp + m] = sw[m] * c;
    Y[int64_t(i) * Kp + K + m] = sw[m] * s;
  }
}

// ------------------------------------------------- FP64 contraction + epilogue
// One (16 kRt) x (16 kRt) tile of A per 256-thread workgroup; thread
// (tx, ty) owns rows ty + 16 r and columns tx + 16 c (r, c < kRt).  K is
// staged through LDS in slices of 16.  A thread reads 2 kRt LDS values per
// k for kRt^2 FMAs: kRt = 4 (64 x 64 tiles) needs 4 bytes of LDS per FMA,
// twice what the LDS can feed the FP64 units; kRt = 8 (128 x 128) halves
// that at a quarter of the workgroups.  Every entry's sum over k runs in
// the same order in both, so they give the same table bit for bit.
constexpr int kSlice = 16;

template <int kRt>
__global__ void __launch_bounds__(256)
table_f64(const double* __restrict__ Y, int n, int Kp,
          const double* __restrict__ pos, MicGeometry g, EwaldScalars es,
          double* __restrict__ A, unsigned long long* __restrict__ pairs) {
  constexpr int kTile = 16 * kRt;
  __shared__ double As[kSlice][kTile + 2];
  __shared__ double Bs[kSlice][kTile + 2];
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  int bi, bj;
  tile_coords_upper(n, kTile, bi, bj);
  const int i0 = bi * kTile, j0 = bj * kTile;
  const bool mirror = bi != bj;

  double acc[kRt][kRt];
#pragma unroll
  for (int r = 0; r < kRt; ++r)
#pragma unroll
    for (int c = 0; c < kRt; ++c) acc[r][c] = 0.0;

  // loader: kTile rows x 16 k = 256 threads x (kTile / 16) doubles per
  // array, as double2 pairs of consecutive k
  constexpr int kPer = kTile / 16;          // doubles per thread and array
  constexpr int kRowsPerPass = 256 * 4 / 16;  // rows per pass of 4 doubles
  auto load_rows = [&](int k0, double (&av)[kPer], double (&bv)[kPer]) {
#pragma unroll
    for (int q = 0; q < kPer / 4; ++q) {
      const int row = (t >> 2) + kRowsPerPass * q;
      const int lk = (t & 3) * 4;
      double2 a01 = make_double2(0.0, 0.0), a23 = a01, b01 = a01, b23 = a01;
      if (i0 + row < n) {
        const double* ya = Y + int64_t(i0 + row) * Kp + k0 + lk;
        a01 = *reinterpret_cast<const double2*>(ya);
        a23 = *reinterpret_cast<const double2*>(ya + 2);
      }
      if (j0 + row < n) {
        const double* yb = Y + int64_t(j0 + row) * Kp + k0 + lk;
        b01 = *reinterpret_cast<const double2*>(yb);
        b23 = *reinterpret_cast<const double2*>(yb + 2);
      }
      av[4 * q] = a01.x;
      av[4 * q + 1] = a01.y;
      av[4 * q + 2] = a23.x;
      av[4 * q + 3] = a23.y;
      bv[4 * q] = b01.x;
      bv[4 * q + 1] = b01.y;
      bv[4 * q + 2] = b23.x;
      bv[4 * q + 3] = b23.y;
    }
  };

  for (int k0 = 0; k0 < Kp; k0 += kSlice) {
    double av[kPer], bv[kPer];
    load_rows(k0, av, bv);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer / 4; ++q) {
      const int row = (t >> 2) + kRowsPerPass * q;
      const int lk = (t & 3) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        As[lk + e][row] = av[4 * q + e];
        Bs[lk + e][row] = bv[4 * q + e];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kSlice; ++kk) {
      double a[kRt], b[kRt];
#pragma unroll
      for (int r = 0; r < kRt; ++r) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < kRt; ++c) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < kRt; ++r)
#pragma unroll
        for (int c = 0; c < kRt; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
  }

  unsigned cnt = 0;
#pragma unroll
  for (int r = 0; r < kRt; ++r) {
    const int i = i0 + ty + 16 * r;
    if (i >= n) continue;
#pragma unroll
    for (int c = 0; c < kRt; ++c) {
      const int j = j0 + tx + 16 * c;
      if (j >= n) continue;
      store_pair(A, n, i, j, mirror, pos, g, es, acc[r][c], cnt);
    }
  }
  count_pairs(cnt, pairs);
}

// ------------------------------------------------------------------ launch
int launch_table_mfma(pycd_ctx* ctx, int precision, const double* pos, int n,
                      const double* kv, const double* sw, int K,
                      const MicGeometry& geo, const EwaldScalars& es,
                      double* out, unsigned long long* pairs,
                      hipEvent_t after_sf, hipEvent_t after_real,
                      bool& real_apart);

int launch_table(pycd_ctx* ctx, int precision, const double* pos, int n,
                 const double* kv, const double* sw, int K,
                 const MicGeometry& geo, const EwaldScalars& es, double* out,
                 unsigned long long* pairs, hipEvent_t after_sf,
                 hipEvent_t after_real, bool& real_apart) {
  real_apart = false;
  if (precision != PYCD_PREC_F64)
    return launch_table_mfma(ctx, precision, pos, n, kv, sw, K, geo, es, out,
                             pairs, after_sf, after_real, real_apart);
  hipStream_t s = ctx->stream;
  const int Kp = ((2 * K + kSlice - 1) / kSlice) * kSlice + (K == 0 ? kSlice : 0);
  DevBuf<double> Y;
  PYCD_HIP(ctx, Y.alloc(size_t(n) * Kp));
  PYCD_HIP(ctx, hipMemsetAsync(Y.ptr, 0, sizeof(double) * size_t(n) * Kp, s));
  if (K > 0) {
    const int64_t total = int64_t(n) * K;
    const int blocks = int(std::min<int64_t>((total + 255) / 256, 65536));
    structure_factor_f64<<<blocks, 256, 0, s>>>(pos, n, kv, sw, K, Kp, Y.ptr);
    PYCD_HIP(ctx, hipGetLastError());
  }
  PYCD_HIP(ctx, hipEventRecord(after_sf, s));
  // (the real-space term is in the epilogue: no time of its own)
  PYCD_HIP(ctx, hipEventRecord(after_real, s));
  // tile variant: 1 = 64 x 64, 2 = 128 x 128; 0 picks 128 from 2048 sites
  int v = ctx->table_variant[PYCD_PREC_F64];
  if (v == 0) v = n >= 2048 ? 2 : 1;
  if (v == 2) {
    const int nb = (n + 127) / 128;
    table_f64<8><<<nb * (nb + 1) / 2, 256, 0, s>>>(Y.ptr, n, Kp, pos, geo, es,
                                                   out, pairs);
  } else {
    const int nb = (n + 63) / 64;
    table_f64<4><<<nb * (nb + 1) / 2, 256, 0, s>>>(Y.ptr, n, Kp, pos, geo, es,
                                                   out, pairs);
  }
  PYCD_HIP(ctx, hipGetLastError());
  // Y is released when this returns; the stream must have consumed it.
  PYCD_HIP(ctx, hipStreamSynchronize(s));
  return PYCD_OK;
}

// -------------------------------------------------------------------- host
int build_ewald_setup(pycd_ctx* ctx, const double* trans, const double* cell,
                      const int32_t* pbc, double alpha, double k_cut,
                      EwaldSetup& st) {
  // Volume and reciprocal lattice from the translational matrix, with the
  // operation order of PyCD/core.py:711-722.
  const double* T = trans;
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  double c12[3], c20[3], c01[3];
  cross(T + 3, T + 6, c12);
  cross(T + 6, T + 0, c20);
  cross(T + 0, T + 3, c01);
  st.volume = std::fabs(T[0] * c12[0] + T[1] * c12[1] + T[2] * c12[2]);
  if (!(st.volume > 0))
    return fail(ctx, PYCD_ERR_ARG, "degenerate translational matrix");
  const double pref = 2 * M_PI / st.volume;
  double B[9];
  for (int d = 0; d < 3; ++d) {
    B[d] = pref * c12[d];
    B[3 + d] = pref * c20[d];
    B[6 + d] = pref * c01[d];
  }
  for (int a = 0; a < 3; ++a) {
    const double len = std::sqrt(B[3 * a] * B[3 * a] + B[3 * a + 1] * B[3 * a + 1] +
                                 B[3 * a + 2] * B[3 * a + 2]);
    st.k_max[a] = int(std::ceil(k_cut / len));
  }
  // Half space in mgrid (C) order: keep v unless -v came earlier, i.e. keep
  // v whose first non-zero component is negative (PyCD/core.py:816-838).
  const double k_cut_2 = k_cut * k_cut;
  const double alpha4 = 4 * alpha;
  st.kv.clear();
  st.sw.clear();
  for (int a = -st.k_max[0]; a <= st.k_max[0]; ++a)
    for (int b = -st.k_max[1]; b <= st.k_max[1]; ++b)
      for (int c = -st.k_max[2]; c <= st.k_max[2]; ++c) {
        const bool keep = a < 0 || (a == 0 && (b < 0 || (b == 0 && c < 0)));
        if (!keep) continue;
        const double k[3] = {a * B[0] + b * B[3] + c * B[6],
                             a * B[1] + b * B[4] + c * B[7],
                             a * B[2] + b * B[5] + c * B[8]};
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        if (!(k2 < k_cut_2)) continue;
        st.kv.insert(st.kv.end(), k, k + 3);
        st.sw.push_back(std::sqrt(pref * std::exp(-k2 / alpha4) / k2));
      }
  st.K = int(st.sw.size());
  // minimum-image geometry
  fill_mic_geometry(cell, pbc, st.geo, /*distance_only=*/true);
  return PYCD_OK;
}

}  // namespace pycd

using namespace pycd;

extern "C" int pycd_pair_table(pycd_ctx* ctx, const double* pos, int64_t n,
                               const double cell[9], const double trans[9],
                               const int32_t pbc[3], double alpha,
                               double r_cut, double k_cut, double eps_r,
                               int32_t precision, double* table_out,
                               pycd_ewald_report* report) {
  if (!ctx) return PYCD_ERR_ARG;
  if (!pos || !cell || !trans || !pbc || !table_out)
    return fail(ctx, PYCD_ERR_ARG, "null argument");
  if (n <= 0 || n > (int64_t(1) << 17))
    return fail(ctx, PYCD_ERR_ARG, "site count %lld outside [1, 131072]",
                (long long)n);
  if (!(alpha > 0) || !(r_cut >= 0) || !(k_cut >= 0) || !(eps_r > 0))
    return fail(ctx, PYCD_ERR_ARG, "invalid Ewald parameters");
  if (precision != PYCD_PREC_F64 && precision != PYCD_PREC_F32 &&
      precision != PYCD_PREC_BF16X3)
    return fail(ctx, PYCD_ERR_ARG, "unknown precision %d", precision);
  DeviceGuard guard(ctx->device);
  EwaldSetup st;
  int rc = build_ewald_setup(ctx, trans, cell, pbc, alpha, k_cut, st);
  if (rc) return rc;
  const int N = int(n);
  const int K = st.K;
  EwaldScalars es;
  es.sqrt_alpha = std::sqrt(alpha);
  es.r_cut = r_cut;
  // fl(r_cut^2) is within half an ulp of the exact square; 2^-50 above it
  // is above the exact square, so r^2 >= r_cut2 gives sqrt(r^2) >= r_cut
  es.r_cut2 = r_cut * r_cut * (1.0 + 0x1p-50);
  es.eps = eps_r;
  es.self_term = -1.0 * std::sqrt(alpha / M_PI) / eps_r;
  es.unit_eps = eps_r == 1.0;

  DevBuf<double> d_pos, d_kv, d_sw, d_A;
  DevBuf<unsigned long long> d_pairs;
  PYCD_HIP(ctx, d_pos.alloc(size_t(N) * 3));
  PYCD_HIP(ctx, d_kv.alloc(size_t(K) * 3 + 3));
  PYCD_HIP(ctx, d_sw.alloc(size_t(K) + 1));
  PYCD_HIP(ctx, d_pairs.alloc(1));
  hipStream_t s = ctx->stream;
  PYCD_HIP(ctx, upload(d_pos.ptr, pos, size_t(N) * 3, s));
  if (K > 0) {
    PYCD_HIP(ctx, hipMemcpyAsync(d_kv.ptr, st.kv.data(), sizeof(double) * 3 * K,
                                 hipMemcpyHostToDevice, s));
    PYCD_HIP(ctx, hipMemcpyAsync(d_sw.ptr, st.sw.data(), sizeof(double) * K,
                                 hipMemcpyHostToDevice, s));
  }
  PYCD_HIP(ctx, hipMemsetAsync(d_pairs.ptr, 0, sizeof(unsigned long long), s));
  double* out = table_out;
  const bool out_dev = is_device_ptr(table_out);
  if (!out_dev) {
    PYCD_HIP(ctx, d_A.alloc(size_t(N) * N));
    out = d_A.ptr;
  }

  // structure factors, [separate real-space kernel], contraction
  hipEvent_t ev[4];
  for (auto& e : ev) PYCD_HIP(ctx, hipEventCreate(&e));
  PYCD_HIP(ctx, hipEventRecord(ev[0], s));
  bool real_apart = false;
  rc = launch_table(ctx, precision, d_pos.ptr, N, d_kv.ptr, d_sw.ptr, K, st.geo,
                    es, out, d_pairs.ptr, ev[1], ev[3], real_apart);
  if (rc) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
  }
  PYCD_HIP(ctx, hipEventRecord(ev[2], s));
  if (!out_dev)
    PYCD_HIP(ctx, hipMemcpyAsync(table_out, out, sizeof(double) * N * size_t(N),
                                 hipMemcpyDeviceToHost, s));
  unsigned long long pairs = 0;
  PYCD_HIP(ctx, hipMemcpyAsync(&pairs, d_pairs.ptr, sizeof(pairs),
                               hipMemcpyDeviceToHost, s));
  PYCD_HIP(ctx, hipStreamSynchronize(s));
  float ms_sf = 0, ms_rest = 0, ms_real = 0;
  (void)hipEventElapsedTime(&ms_sf, ev[0], ev[1]);
  (void)hipEventElapsedTime(&ms_rest, ev[1], ev[2]);
  (void)hipEventElapsedTime(&ms_real, ev[1], ev[3]);
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (report) {
    report->n_kvec = K;
    report->n_real_pairs = int64_t(pairs);
    for (int a = 0; a < 3; ++a) report->k_max[a] = st.k_max[a];
    report->precision = precision;
    report->volume = st.volume;
    report->ms_sf = ms_sf;
    report->ms_real = real_apart ? ms_real : 0.0;
    report->ms_recip = ms_rest;
    report->ms_total = double(ms_sf) + ms_rest;
  }
  return PYCD_OK;
}

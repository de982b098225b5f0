// Windowed-origin squared displacement of Analysis.compute_msd on gfx950.
//
//   sd[t, tau, c] = mean_{t0 < P - tau} |r(t0 + tau) - r(t0)|^2
// for every trajectory t, lag tau in [1, M) and carrier c
// (PyCD/core.py:2996-3007); sd[t, 0, c] = 0.
//
// One workgroup per (trajectory, carrier, block of 8 lags).  The block's
// carrier path is read from L2 once per lag; each thread accumulates a strip
// of origins for all 8 lags at once so that r(t0) is loaded once per 8 lags.
#include "common.hpp"

namespace pycd {

constexpr int kLagBlock = 8;

__global__ void __launch_bounds__(256)
msd_kernel(const double* __restrict__ pos, int64_t P, int nc, int64_t M,
           double* __restrict__ sd) {
  const int64_t nlb = (M + kLagBlock - 1) / kLagBlock;
  const int64_t b = blockIdx.x;
  const int64_t lb = b % nlb;
  const int64_t rest = b / nlb;
  const int c = int(rest % nc);
  const int64_t traj = rest / nc;
  const int64_t tau0 = lb * kLagBlock;
  const double* base = pos + (traj * P * nc + c) * 3;
  const int64_t stride = int64_t(nc) * 3;
  double acc[kLagBlock];
#pragma unroll
  for (int l = 0; l < kLagBlock; ++l) acc[l] = 0.0;
  for (int64_t t0 = threadIdx.x; t0 < P; t0 += blockDim.x) {
    const double x0 = base[t0 * stride];
    const double y0 = base[t0 * stride + 1];
    const double z0 = base[t0 * stride + 2];
#pragma unroll
    for (int l = 0; l < kLagBlock; ++l) {
      const int64_t tau = tau0 + l;
      if (tau >= 1 && tau < M && t0 + tau < P) {
        const double* q = base + (t0 + tau) * stride;
        const double dx = q[0] - x0, dy = q[1] - y0, dz = q[2] - z0;
        acc[l] += dx * dx + dy * dy + dz * dz;
      }
    }
  }
  __shared__ double part[kLagBlock][256];
#pragma unroll
  for (int l = 0; l < kLagBlock; ++l) part[l][threadIdx.x] = acc[l];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
#pragma unroll
      for (int l = 0; l < kLagBlock; ++l)
        part[l][threadIdx.x] += part[l][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kLagBlock) {
    const int64_t tau = tau0 + threadIdx.x;
    if (tau < M) {
      const double v = (tau >= 1 && tau < P)
                           ? part[threadIdx.x][0] / double(P - tau)
                           : 0.0;
      sd[(traj * M + tau) * nc + c] = v;
    }
  }
}

}  // namespace pycd

using namespace pycd;

extern "C" int pycd_msd(pycd_ctx* ctx, const double* pos, int64_t n_traj,
                        int64_t n_path, int64_t n_carriers, int64_t n_lags,
                        double* sd) {
  if (!ctx) return PYCD_ERR_ARG;
  if (!pos || !sd || n_traj <= 0 || n_path <= 0 || n_carriers <= 0 ||
      n_lags <= 0)
    return fail(ctx, PYCD_ERR_ARG, "bad MSD arguments");
  if (n_lags > n_path)
    return fail(ctx, PYCD_ERR_ARG, "n_lags %lld exceeds n_path %lld",
                (long long)n_lags, (long long)n_path);
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  const size_t npos = size_t(n_traj) * n_path * n_carriers * 3;
  const size_t nsd = size_t(n_traj) * n_lags * n_carriers;
  DevBuf<double> dpos, dsd;
  const double* p = pos;
  double* o = sd;
  if (!is_device_ptr(pos)) {
    PYCD_HIP(ctx, dpos.alloc(npos));
    PYCD_HIP(ctx, hipMemcpyAsync(dpos.ptr, pos, 8 * npos, hipMemcpyHostToDevice,
                                 st));
    p = dpos.ptr;
  }
  const bool out_dev = is_device_ptr(sd);
  if (!out_dev) {
    PYCD_HIP(ctx, dsd.alloc(nsd));
    o = dsd.ptr;
  }
  const int64_t nlb = (n_lags + kLagBlock - 1) / kLagBlock;
  const int64_t blocks = n_traj * n_carriers * nlb;
  if (blocks > (int64_t(1) << 31) - 1)
    return fail(ctx, PYCD_ERR_ARG, "MSD problem too large for one launch");
  msd_kernel<<<unsigned(blocks), 256, 0, st>>>(p, n_path, int(n_carriers),
                                                n_lags, o);
  PYCD_HIP(ctx, hipGetLastError());
  if (!out_dev)
    PYCD_HIP(ctx, hipMemcpyAsync(sd, o, 8 * nsd, hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipStreamSynchronize(st));
  return PYCD_OK;
}

// Minimum-image vectors and hop neighbour lists on gfx950.
//
//   pycd_min_image_vectors  <- Neighbors.get_pairwise_min_image_vector_data
//                              (PyCD/core.py:571-594): N x N x 3 vectors
//                              pos[j] - pos[i] under the minimum image
//   pycd_hop_neighbors      <- Neighbors.hop_neighbor_sites /
//                              generate_neighbor_list (PyCD/core.py:485-668):
//                              per centre and shell, the targets with
//                              lo < |r| <= hi in ascending target order
//
// The minimum image follows PyCD/core.py:304-361: fractional coordinates,
// round-half-even base offset on periodic axes, then the 26 neighbouring
// offsets in (dx, dy, dz) loop order, a candidate replacing the best only if
// strictly shorter.
#include "common.hpp"
#include "ewald_common.hpp"

namespace pycd {

__device__ inline double mic_vec(double dx, double dy, double dz,
                                 const MicGeometry& g, double* v) {
  double f0 = dx * g.inv[0] + dy * g.inv[3] + dz * g.inv[6];
  double f1 = dx * g.inv[1] + dy * g.inv[4] + dz * g.inv[7];
  double f2 = dx * g.inv[2] + dy * g.inv[5] + dz * g.inv[8];
  if (g.pbc[0]) f0 -= rint(f0);
  if (g.pbc[1]) f1 -= rint(f1);
  if (g.pbc[2]) f2 -= rint(f2);
  double bx = f0 * g.cell[0] + f1 * g.cell[3] + f2 * g.cell[6];
  double by = f0 * g.cell[1] + f1 * g.cell[4] + f2 * g.cell[7];
  double bz = f0 * g.cell[2] + f1 * g.cell[5] + f2 * g.cell[8];
  double best = bx * bx + by * by + bz * bz;
  double vx = bx, vy = by, vz = bz;
  for (int m = 0; m < g.n_delta; ++m) {
    const double cx = bx + g.dS[m][0];
    const double cy = by + g.dS[m][1];
    const double cz = bz + g.dS[m][2];
    const double r2 = cx * cx + cy * cy + cz * cz;
    if (r2 < best) {
      best = r2;
      vx = cx;
      vy = cy;
      vz = cz;
    }
  }
  v[0] = vx;
  v[1] = vy;
  v[2] = vz;
  return sqrt(best);
}

__global__ void __launch_bounds__(256)
min_image_kernel(const double* __restrict__ pos, int64_t n, MicGeometry g,
                 double* __restrict__ out) {
  const int64_t total = n * n;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < total; idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx / n, j = idx - i * n;
    double v[3];
    mic_vec(pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
            pos[3 * j + 2] - pos[3 * i + 2], g, v);
    out[3 * idx] = v[0];
    out[3 * idx + 1] = v[1];
    out[3 * idx + 2] = v[2];
  }
}

constexpr int kMaxShells = 16;

struct ShellWindows {
  double lo[kMaxShells];
  double hi[kMaxShells];
  int n;
};

// One wavefront per centre; lanes sweep the targets 64 at a time in
// ascending order, and a ballot per shell gives each hit its rank.
__global__ void __launch_bounds__(64)
hop_neighbor_kernel(const double* __restrict__ pos,
                    const int64_t* __restrict__ centers, int64_t n_centers,
                    const int64_t* __restrict__ targets, int64_t n_targets,
                    MicGeometry g, ShellWindows w, int max_per,
                    int32_t* __restrict__ count, int32_t* __restrict__ index,
                    double* __restrict__ vec) {
  const int lane = threadIdx.x;
  const int64_t c = blockIdx.x;
  if (c >= n_centers) return;
  const int64_t ci = centers[c];
  const double px = pos[3 * ci], py = pos[3 * ci + 1], pz = pos[3 * ci + 2];
  int running[kMaxShells];
#pragma unroll
  for (int s = 0; s < kMaxShells; ++s) running[s] = 0;
  const unsigned long long below =
      lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int64_t j0 = 0; j0 < n_targets; j0 += 64) {
    const int64_t j = j0 + lane;
    double v[3] = {0.0, 0.0, 0.0};
    double r = -1.0;
    if (j < n_targets) {
      const int64_t tj = targets[j];
      r = mic_vec(pos[3 * tj] - px, pos[3 * tj + 1] - py, pos[3 * tj + 2] - pz,
                  g, v);
    }
#pragma unroll
    for (int s = 0; s < kMaxShells; ++s) {
      if (s >= w.n) break;
      const bool in = j < n_targets && w.lo[s] < r && r <= w.hi[s];
      const unsigned long long mask = __ballot(in);
      if (in) {
        const int rank = running[s] + __popcll(mask & below);
        if (rank < max_per) {
          const int64_t o = (c * w.n + s) * max_per + rank;
          index[o] = int32_t(j);
          vec[3 * o] = v[0];
          vec[3 * o + 1] = v[1];
          vec[3 * o + 2] = v[2];
        }
      }
      running[s] += __popcll(mask);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < kMaxShells; ++s)
      if (s < w.n) count[c * w.n + s] = running[s];
  }
}

}  // namespace pycd

using namespace pycd;

extern "C" int pycd_min_image_vectors(pycd_ctx* ctx, const double* pos,
                                      int64_t n, const double cell[9],
                                      const int32_t pbc[3], double* out) {
  if (!ctx) return PYCD_ERR_ARG;
  if (!pos || !cell || !pbc || !out || n <= 0 || n > (int64_t(1) << 17))
    return fail(ctx, PYCD_ERR_ARG, "bad minimum-image arguments");
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  MicGeometry g;
  fill_mic_geometry(cell, pbc, g);
  DevBuf<double> dpos, dout;
  PYCD_HIP(ctx, dpos.alloc(size_t(n) * 3));
  PYCD_HIP(ctx, upload(dpos.ptr, pos, size_t(n) * 3, st));
  double* o = out;
  const bool out_dev = is_device_ptr(out);
  if (!out_dev) {
    PYCD_HIP(ctx, dout.alloc(size_t(n) * n * 3));
    o = dout.ptr;
  }
  const int64_t total = n * n;
  const int blocks = int(std::min<int64_t>((total + 255) / 256, 1 << 20));
  min_image_kernel<<<blocks, 256, 0, st>>>(dpos.ptr, n, g, o);
  PYCD_HIP(ctx, hipGetLastError());
  if (!out_dev)
    PYCD_HIP(ctx, hipMemcpyAsync(out, o, 8 * 3 * size_t(total),
                                 hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipStreamSynchronize(st));
  return PYCD_OK;
}

extern "C" int pycd_hop_neighbors(pycd_ctx* ctx, const double* pos, int64_t n,
                                  const int64_t* centers, int64_t n_centers,
                                  const int64_t* targets, int64_t n_targets,
                                  const double cell[9], const int32_t pbc[3],
                                  const double* lo, const double* hi,
                                  int32_t n_shells, int32_t max_per_shell,
                                  int32_t* count, int32_t* index,
                                  double* vec) {
  if (!ctx) return PYCD_ERR_ARG;
  if (!pos || !centers || !targets || !cell || !pbc || !lo || !hi || !count ||
      !index || !vec || n <= 0 || n_centers <= 0 || n_targets <= 0 ||
      n_shells <= 0 || n_shells > kMaxShells || max_per_shell <= 0 ||
      n_centers > (int64_t(1) << 24) || n_targets > (int64_t(1) << 24))
    return fail(ctx, PYCD_ERR_ARG, "bad neighbour-list arguments");
  for (int64_t k = 0; k < n_centers; ++k)
    if (centers[k] < 0 || centers[k] >= n)
      return fail(ctx, PYCD_ERR_ARG, "centre index out of range");
  for (int64_t k = 0; k < n_targets; ++k)
    if (targets[k] < 0 || targets[k] >= n)
      return fail(ctx, PYCD_ERR_ARG, "target index out of range");
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  MicGeometry g;
  fill_mic_geometry(cell, pbc, g);
  ShellWindows w;
  w.n = n_shells;
  for (int s = 0; s < kMaxShells; ++s) {
    w.lo[s] = s < n_shells ? lo[s] : 0.0;
    w.hi[s] = s < n_shells ? hi[s] : -1.0;
  }
  const size_t nout = size_t(n_centers) * n_shells * max_per_shell;
  DevBuf<double> dpos, dvec;
  DevBuf<int64_t> dc, dt;
  DevBuf<int32_t> dcount, dindex;
  PYCD_HIP(ctx, dpos.alloc(size_t(n) * 3));
  PYCD_HIP(ctx, dc.alloc(n_centers));
  PYCD_HIP(ctx, dt.alloc(n_targets));
  PYCD_HIP(ctx, dcount.alloc(size_t(n_centers) * n_shells));
  PYCD_HIP(ctx, dindex.alloc(nout));
  PYCD_HIP(ctx, dvec.alloc(nout * 3));
  PYCD_HIP(ctx, upload(dpos.ptr, pos, size_t(n) * 3, st));
  PYCD_HIP(ctx, hipMemcpyAsync(dc.ptr, centers, 8 * n_centers,
                               hipMemcpyHostToDevice, st));
  PYCD_HIP(ctx, hipMemcpyAsync(dt.ptr, targets, 8 * n_targets,
                               hipMemcpyHostToDevice, st));
  hop_neighbor_kernel<<<unsigned(n_centers), 64, 0, st>>>(
      dpos.ptr, dc.ptr, n_centers, dt.ptr, n_targets, g, w, max_per_shell,
      dcount.ptr, dindex.ptr, dvec.ptr);
  PYCD_HIP(ctx, hipGetLastError());
  PYCD_HIP(ctx, hipMemcpyAsync(count, dcount.ptr,
                               4 * size_t(n_centers) * n_shells,
                               hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipMemcpyAsync(index, dindex.ptr, 4 * nout,
                               hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipMemcpyAsync(vec, dvec.ptr, 8 * 3 * nout,
                               hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipStreamSynchronize(st));
  for (size_t k = 0; k < size_t(n_centers) * n_shells; ++k)
    if (count[k] > max_per_shell)
      return fail(ctx, PYCD_ERR_ARG,
                  "a shell holds %d targets (> max_per_shell %d); counts are "
                  "returned",
                  count[k], max_per_shell);
  return PYCD_OK;
}

// Pieces shared by the FP64 and MFMA Ewald table kernels.
#pragma once

#include "common.hpp"

#include <cmath>
#include <vector>

namespace pycd {

// Minimum-image geometry: frac = d S^-1, wrap periodic axes with round-half-
// even, then try the neighbouring images, keeping a strictly shorter one
// (PyCD/core.py:304-361).  dS holds delta . S for the candidate offsets: the
// 26 neighbours, less those that can never be shorter than the wrapped
// vector for this cell (fill_mic_geometry).
struct MicGeometry {
  double inv[9];
  double cell[9];
  double dS[26][3];
  int n_delta;
  int pbc[3];
};

struct EwaldScalars {
  double sqrt_alpha;
  double r_cut;
  double r_cut2;     // a few ulp above r_cut^2: r^2 >= r_cut2 gives r >= r_cut
  double eps;
  double self_term;  // -sqrt(alpha/pi)/eps
  int unit_eps;      // eps == 1, where x / eps == x
};

struct EwaldSetup {
  double volume = 0;
  int k_max[3] = {0, 0, 0};
  int K = 0;
  std::vector<double> kv;  // K x 3 reciprocal vectors (bohr^-1)
  std::vector<double> sw;  // sqrt of (2pi/V) exp(-k^2/(4 alpha)) / k^2
  MicGeometry geo;
};

// distance_only: the users need the minimum distance, not which image
// gives it (the Ewald table), so offsets that can only tie are dropped too.
inline void fill_mic_geometry(const double* cell, const int32_t* pbc,
                              MicGeometry& g, bool distance_only = false) {
  for (int k = 0; k < 9; ++k) g.cell[k] = cell[k];
  const double* m = cell;
  const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) -
                     m[1] * (m[3] * m[8] - m[5] * m[6]) +
                     m[2] * (m[3] * m[7] - m[4] * m[6]);
  g.inv[0] = (m[4] * m[8] - m[5] * m[7]) / det;
  g.inv[1] = (m[2] * m[7] - m[1] * m[8]) / det;
  g.inv[2] = (m[1] * m[5] - m[2] * m[4]) / det;
  g.inv[3] = (m[5] * m[6] - m[3] * m[8]) / det;
  g.inv[4] = (m[0] * m[8] - m[2] * m[6]) / det;
  g.inv[5] = (m[2] * m[3] - m[0] * m[5]) / det;
  g.inv[6] = (m[3] * m[7] - m[4] * m[6]) / det;
  g.inv[7] = (m[1] * m[6] - m[0] * m[7]) / det;
  g.inv[8] = (m[0] * m[4] - m[1] * m[3]) / det;
  for (int a = 0; a < 3; ++a) g.pbc[a] = pbc[a] ? 1 : 0;
  const bool all_pbc = g.pbc[0] && g.pbc[1] && g.pbc[2];
  g.n_delta = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int d[3] = {g.pbc[0] ? dx : 0, g.pbc[1] ? dy : 0,
                          g.pbc[2] ? dz : 0};
        if ((dx == 0 && dy == 0 && dz == 0) || (!d[0] && !d[1] && !d[2]))
          continue;
        double v[3];
        for (int c = 0; c < 3; ++c)
          v[c] = d[0] * cell[c] + d[1] * cell[3 + c] + d[2] * cell[6 + c];
        if (all_pbc) {
          // b + v is shorter than the wrapped b = f S (|f_k| <= 1/2) only if
          // b.v < -|v|^2 / 2, and min over f of b.v = -1/2 sum_k |S_k . v|:
          // an offset with sum_k |S_k . v| <= |v|^2 can at best tie (on the
          // boundary of the wrapped cell), so dropping it leaves the minimum
          // distance unchanged (to rounding where the reference's strict
          // comparison would see an exact tie).  Hexagonal and monoclinic
          // supercells keep 12 of the 26 offsets here, orthogonal ones none.
          double reach = 0.0;
          for (int k = 0; k < 3; ++k)
            reach += std::fabs(cell[3 * k] * v[0] + cell[3 * k + 1] * v[1] +
                               cell[3 * k + 2] * v[2]);
          const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
          if (reach <= v2 * (1.0 + 1e-12)) continue;
        }
        for (int c = 0; c < 3; ++c) g.dS[g.n_delta][c] = v[c];
        ++g.n_delta;
      }
  if (all_pbc && distance_only) {
    // An offset v is also never needed when another kept offset u is at
    // least as short for every wrapped b: |b + v|^2 - |b + u|^2 =
    // 2 b.(v - u) + |v|^2 - |u|^2, whose minimum over b is
    // |v|^2 - |u|^2 - sum_k |S_k . (v - u)|.  (With c normal to the a-b
    // plane, every offset with a c part loses to its in-plane part:
    // hexagonal and monoclinic supercells keep 4 offsets of 12.)  A
    // dropped offset only ties, on the wrapped cell's boundary, with the
    // offset that beats it: the minimum distance stays the same, the image
    // that gives it may not.
    bool dropped = true;
    while (dropped) {
      dropped = false;
      for (int a = 0; a < g.n_delta && !dropped; ++a)
        for (int b = 0; b < g.n_delta && !dropped; ++b) {
          if (a == b) continue;
          const double* v = g.dS[a];
          const double* u = g.dS[b];
          const double w[3] = {v[0] - u[0], v[1] - u[1], v[2] - u[2]};
          double reach = 0.0;
          for (int k = 0; k < 3; ++k)
            reach += std::fabs(cell[3 * k] * w[0] + cell[3 * k + 1] * w[1] +
                               cell[3 * k + 2] * w[2]);
          const double gain = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) -
                              (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
          if (reach <= gain * (1.0 + 1e-12)) {
            for (int m = a; m + 1 < g.n_delta; ++m)
              for (int c = 0; c < 3; ++c) g.dS[m][c] = g.dS[m + 1][c];
            --g.n_delta;
            dropped = true;
          }
        }
    }
  }
  for (int k = g.n_delta; k < 26; ++k)
    g.dS[k][0] = g.dS[k][1] = g.dS[k][2] = 0.0;
}

__device__ __forceinline__ double mic_r2(double dx, double dy, double dz,
                                const MicGeometry& g) {
  double f0 = dx * g.inv[0] + dy * g.inv[3] + dz * g.inv[6];
  double f1 = dx * g.inv[1] + dy * g.inv[4] + dz * g.inv[7];
  double f2 = dx * g.inv[2] + dy * g.inv[5] + dz * g.inv[8];
  if (g.pbc[0]) f0 -= rint(f0);
  if (g.pbc[1]) f1 -= rint(f1);
  if (g.pbc[2]) f2 -= rint(f2);
  const double bx = f0 * g.cell[0] + f1 * g.cell[3] + f2 * g.cell[6];
  const double by = f0 * g.cell[1] + f1 * g.cell[4] + f2 * g.cell[7];
  const double bz = f0 * g.cell[2] + f1 * g.cell[5] + f2 * g.cell[8];
  double best = bx * bx + by * by + bz * bz;
  for (int m = 0; m < g.n_delta; ++m) {
    const double cx = bx + g.dS[m][0];
    const double cy = by + g.dS[m][1];
    const double cz = bz + g.dS[m][2];
    const double r2 = cx * cx + cy * cy + cz * cz;
    best = r2 < best ? r2 : best;
  }
  return best;
}

// Real-space part of entry (i, j): erfc(sqrt(alpha) r) / (2 r) / eps within
// r_cut (the diagonal without the 1/r), else 0; counts the pair in `cnt`.
// The square root only near the cutoff and no division by a unit eps: the
// same numbers as r = sqrt(r^2), r < r_cut and x / eps throughout.
// (dx, dy, dz) = pos[j] - pos[i].
__device__ __forceinline__ double real_part_d(int i, int j, double dx,
                                              double dy, double dz,
                                              const MicGeometry& g,
                                              const EwaldScalars& es,
                                              unsigned& cnt) {
  const double r2 = mic_r2(dx, dy, dz, g);
  double real = 0.0;
  if (r2 < es.r_cut2) {
    const double r = sqrt(r2);
    if (r < es.r_cut) {
      ++cnt;
      double v = erfc(es.sqrt_alpha * r) / 2;
      if (i != j) v /= r;
      real = es.unit_eps ? v : v / es.eps;
    }
  }
  return real;
}

__device__ __forceinline__ double real_part(int i, int j,
                                            const double* __restrict__ pos,
                                            const MicGeometry& g,
                                            const EwaldScalars& es,
                                            unsigned& cnt) {
  return real_part_d(i, j, pos[3 * j] - pos[3 * i],
                     pos[3 * j + 1] - pos[3 * i + 1],
                     pos[3 * j + 2] - pos[3 * i + 2], g, es, cnt);
}

// Entry (i, j) from its real-space part and the reciprocal accumulator
// `acc` (= sum over the half space of w cos, i.e. (Y Y^T)[i,j]).
__device__ __forceinline__ double finish_entry(int i, int j,
                                               const EwaldScalars& es,
                                               double real, double acc) {
  double out = real + (es.unit_eps ? 2.0 * acc : (2.0 * acc) / es.eps);
  if (i == j) out += es.self_term;
  return out;
}

__device__ __forceinline__ double assemble_entry(int i, int j,
                                        const double* __restrict__ pos,
                                        const MicGeometry& g,
                                        const EwaldScalars& es, double acc,
                                        unsigned& cnt) {
  return finish_entry(i, j, es, real_part(i, j, pos, g, es, cnt), acc);
}

__device__ inline void count_pairs(unsigned cnt,
                                   unsigned long long* __restrict__ pairs) {
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0 && cnt)
    atomicAdd(pairs, static_cast<unsigned long long>(cnt));
}

// The table is symmetric (Y Y^T, and the real-space and self terms), so
// only tiles with bi <= bj are launched: nb (nb + 1) / 2 workgroups.
// Consecutive workgroups are dealt round-robin to the 8 XCDs; each XCD
// walks a contiguous run of tiles.  The run goes through groups of 8 tile
// rows column by column (each column's rows bi <= bj of the group), so that
// a column block of Y serves up to 8 consecutive tiles and the group's row
// blocks stay in cache along the walk.
// Tile of position w in that order (host and device).
__host__ __device__ inline void tile_coords_upper_at(int n, int tile, int w,
                                                     int& bi, int& bj) {
  const int nb = (n + tile - 1) / tile;
  const int T = nb * (nb + 1) / 2;
  const int full = (T / 8) * 8;
  if (w < full) {
    const int per = full / 8;
    w = (w % 8) * per + (w / 8);
  }
  constexpr int G = 8;
  for (int g0 = 0; g0 < nb; g0 += G) {
    const int gs = G < nb - g0 ? G : nb - g0;
    const int head = gs * (gs + 1) / 2;  // columns g0 .. g0 + gs - 1
    const int count = head + (nb - g0 - gs) * gs;
    if (w >= count) {
      w -= count;
      continue;
    }
    if (w < head) {
      int c = 0;  // column g0 + c holds rows g0 .. g0 + c
      while (w >= c + 1) {
        w -= c + 1;
        ++c;
      }
      bj = g0 + c;
      bi = g0 + w;
    } else {
      w -= head;
      bj = g0 + gs + w / gs;
      bi = g0 + w % gs;
    }
    return;
  }
  bi = bj = 0;  // not reached: w < T
}

__device__ inline void tile_coords_upper(int n, int tile, int& bi, int& bj) {
  tile_coords_upper_at(n, tile, int(blockIdx.x), bi, bj);
}

// Entry (i, j) of a tile and, off the diagonal block, its mirror (j, i),
// which is the same number: the mirror's reciprocal sum is the same
// accumulation (the same products in the same order), and its real-space
// part comes from pos[i] - pos[j] = -(pos[j] - pos[i]) exactly, whose
// fractional coordinates, round-half-even wrap and candidate images (the
// kept offsets come in +/- pairs) are the negatives of those of (i, j), so
// the squared minimum-image distance and everything after it are equal.
// The FP64 epilogue is therefore formed once per pair.
__device__ __forceinline__ void store_pair(double* __restrict__ A, int n, int i, int j,
                                  bool mirror, const double* __restrict__ pos,
                                  const MicGeometry& g, const EwaldScalars& es,
                                  double acc, unsigned& cnt) {
  const unsigned before = cnt;
  const double v = assemble_entry(i, j, pos, g, es, acc, cnt);
  A[int64_t(i) * n + j] = v;
  if (mirror) {
    A[int64_t(j) * n + i] = v;
    cnt += cnt - before;  // ordered pairs within the cutoff: (j, i) too
  }
}

int launch_table(pycd_ctx* ctx, int precision, const double* pos, int n,
                 const double* kv, const double* sw, int K,
                 const MicGeometry& geo, const EwaldScalars& es, double* out,
                 unsigned long long* pairs, hipEvent_t after_sf,
                 hipEvent_t after_real, bool& real_apart);

}  // namespace pycd

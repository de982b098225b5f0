// Packed site codes of a cell-periodic supercell, shared by the batched and
// plain KMC kernels.  Sites are indexed as PyCD indexes them,
// ((cx ny + cy) nz + cz) n_unit + sub (PyCD/core.py:201-235).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pycd {

// ---- packed sites of a cell-periodic family
// A site is cx | cz << 8 | cy << 16 | sub << 24: masking with 0x00ff00ff
// leaves (cx, cy) as two 16-bit lanes, so the wrapped cell difference of x
// and y takes packed 16-bit integer instructions, and the column a
// difference names is one 16-bit dot product.
struct Geo {
  int nx, ny, nz, nu;
};

__host__ __device__ inline uint32_t site_pack(uint32_t i, const Geo& g) {
  const uint32_t sub = i % uint32_t(g.nu);
  uint32_t c = i / uint32_t(g.nu);
  const uint32_t cz = c % uint32_t(g.nz);
  c /= uint32_t(g.nz);
  const uint32_t cy = c % uint32_t(g.ny);
  const uint32_t cx = c / uint32_t(g.ny);
  return cx | (cz << 8) | (cy << 16) | (sub << 24);
}

__host__ __device__ inline uint32_t site_index(uint32_t p, const Geo& g) {
  const uint32_t cx = p & 0xffu, cz = (p >> 8) & 0xffu,
                 cy = (p >> 16) & 0xffu;
  return ((cx * uint32_t(g.ny) + cy) * uint32_t(g.nz) + cz) * uint32_t(g.nu) +
         (p >> 24);
}

// Site s moved by the packed cell offset d (components in [0, n)), landing
// on the sublattice that d names.
__host__ __device__ inline uint32_t site_add(uint32_t s, uint32_t d,
                                             const Geo& g) {
  int x = int(s & 0xffu) + int(d & 0xffu);
  x -= x >= g.nx ? g.nx : 0;
  int w = int((s >> 8) & 0xffu) + int((d >> 8) & 0xffu);
  w -= w >= g.nz ? g.nz : 0;
  int y = int((s >> 16) & 0xffu) + int((d >> 16) & 0xffu);
  y -= y >= g.ny ? g.ny : 0;
  return uint32_t(x) | (uint32_t(w) << 8) | (uint32_t(y) << 16) |
         (d & 0xff000000u);
}

// A table of a periodic supercell is used through the rows of cell 0 only
// if every entry equals its cell-0 translate to this fraction of the
// table's largest entry (translates built by the same arithmetic differ by
// ~1e-16 of it).  NaN anywhere fails the check.
constexpr double kTranslationTol = 1e-14;

// max(a, b) that keeps a NaN of either side (fmax drops it).
__host__ __device__ inline double nan_max(double a, double b) {
  return (b != b || b > a) ? b : a;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Per-row part of a column lookup: the compact row of a's sublattice, and
// (n - a) per axis for the wrapped difference.
struct RowKey {
  u16x2 kxy;     // (nx - ax, ny - ay)
  uint32_t kz;   // nz - az
  uint32_t sub;  // a's sublattice
};

// Launch-constant packed sizes and strides (N < 2^24 and ny nz nu < 2^16).
struct GeoPk {
  u16x2 nxy;     // (nx, ny)
  u16x2 sxy;     // (ny nz nu, nz nu)
  uint32_t nz, nu;
};

__host__ __device__ inline GeoPk geo_pk(const Geo& g) {
  GeoPk p;
  p.nxy = u16x2{uint16_t(g.nx), uint16_t(g.ny)};
  p.sxy = u16x2{uint16_t(g.ny * g.nz * g.nu), uint16_t(g.nz * g.nu)};
  p.nz = uint32_t(g.nz);
  p.nu = uint32_t(g.nu);
  return p;
}

__device__ inline RowKey row_key(uint32_t a, const GeoPk& p) {
  RowKey k;
  k.kxy = p.nxy - __builtin_bit_cast(u16x2, a & 0x00ff00ffu);
  k.kz = p.nz - ((a >> 8) & 0xffu);
  k.sub = a >> 24;
  return k;
}

// Column of site b in the compact row of key's site: b moved by -cell(a),
// as a site index (the wrapped difference is min(t, t - n) in unsigned
// arithmetic, t = b + n - a in [1, 2n)).
__device__ inline uint32_t rel_col(const RowKey& k, uint32_t b,
                                   const GeoPk& p) {
  const u16x2 t = __builtin_bit_cast(u16x2, b & 0x00ff00ffu) + k.kxy;
  const u16x2 dxy = __builtin_elementwise_min(t, t - p.nxy);
  const uint32_t tz = ((b >> 8) & 0xffu) + k.kz;
  const uint32_t dz = min(tz, tz - p.nz);
  return __builtin_amdgcn_udot2(dxy, p.sxy, __umul24(dz, p.nu) + (b >> 24),
                                false);
}

__device__ inline uint32_t rel_col(uint32_t a, uint32_t b, const GeoPk& p) {
  return rel_col(row_key(a, p), b, p);
}

}  // namespace pycd

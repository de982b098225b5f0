// Plain (one-trajectory) KMC loop of PyCD's Run.do_kmc_steps on gfx950.
//
// Per hop (PyCD/core.py:2787-2861, rates from :1989-2050):
//   dG0_p = 2 q_s [ q.(A[n,:] - A[s,:]) + q_s (A[s,s] - A[s,n]) ] + E[n] - E[s]
//   dGs_p = ((lambda + dG0)^2 / (4 lambda) - V_ab) - 0.5 F.d
//   k_p   = nu * e**(-dGs / kT)
//   select the first p with cumsum(k / k_tot)_p > u1; t -= ln(u2) / k_tot
//
// The N-long dot product is never formed.  With q = q_lattice + sum_c q_c
// e_{occ_c},  q.A[x,:] = phi_L[x] + psi[x],  where phi_L = A q_lattice is
// computed once per table and psi[x] = sum_c q_c A[x, occ_c] is gathered
// each hop for the sites x that matter: every carrier's site and its z hop
// targets.  That is n_c (1 + z) x n_c table reads per hop instead of
// n_c z x N, with the same value of q.(A[n,:] - A[s,:]) up to rounding.
//
// One persistent 256-thread workgroup runs the trajectory (the hop chain is
// serial).  Phases per hop, separated by workgroup barriers:
//   A  psi gathers                                    (all threads)
//   B  dG0, Marcus/field rate per process             (all threads)
//   C  k_tot, wavefront prefix scan of k/k_tot,
//      selection, clock, state update                 (wavefront 0)
//   D  path-step rows written for [start, end)        (all threads)
#include "common.hpp"
#include "packed_sites.hpp"
#include "wave.hpp"

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <cmath>

namespace pycd {

struct KmcScalars {
  double t;
  double energy;
  int64_t start;
  int64_t hops;
  int32_t finished;
  int32_t no_select;  // hops where rounding left no cum > u1 (last taken)
};

struct KmcDev {
  const double* A;
  // periodic supercell (compact instance): the table rows of cell 0, the
  // packed code of every site and of every slot's target
  const double* Ac;
  const uint32_t* pk;
  const uint32_t* nbr_pk;
  GeoPk gp;
  const double* phiL;
  const double* dphi;  // q_L . (A[n, :] - A[s, :]) per (site, slot)
  const int32_t* nbr;
  const double* vec;
  const double* lam;
  const double* vab;
  const double* tself;
  const double* E;
  const double* cq;
  int64_t N;
  int32_t nc, z, max_z, field_on;
  double field[3];
  double nu, kT, dt;
  double nkt;  // -1 / kT
  int64_t P;
  // state
  int32_t* occ;
  double* disp;
  double* base;
  double* drift;
  double* rows;
  double* epath;
  double* dpath;
  KmcScalars* sc;
};

constexpr int kThreads = 256;

// log(u2) of every uniform pair ahead of the hop loop (0 where u2 = 0, a
// value the loop does not use), so that no hop waits for an FP64 log.
__global__ void __launch_bounds__(256)
log_u2_kernel(const double* __restrict__ U, int64_t n,
              double* __restrict__ lu) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double u2 = U[2 * i + 1];
    lu[i] = u2 > 0.0 ? log(u2) : 0.0;
  }
}

template <bool kCompact>
__global__ void __launch_bounds__(kThreads)
kmc_plain_kernel(KmcDev d, const double* __restrict__ U,
                 const double* __restrict__ LU, int64_t n_pairs,
                 int64_t max_hops, double* __restrict__ hop_time,
                 int32_t* __restrict__ hop_proc, int32_t* __restrict__ status) {
  extern __shared__ double smem[];
  const int nc = d.nc, z = d.z, R = nc * (1 + z), nproc = nc * z;
  double* psi = smem;                 // R
  double* kk = psi + R;               // nproc
  double* dg = kk + nproc;            // nproc
  double* disp = dg + nproc;          // 3 nc
  double* base = disp + 3 * nc;       // 3 nc
  double* drift = base + 3 * nc;      // 3 nc
  // per process p = c z + j (carrier c, slot j of its site): the values of
  // its rate that depend only on the site, refreshed when c hops
  double* pdphi = drift + 3 * nc;     // q_L . (A[n, :] - A[s, :])
  double* ptq = pdphi + nproc;        // q_c tself[s, j]
  double* pdE = ptq + nproc;          // E[n] - E[s]
  double* plam = pdE + nproc;         // lambda
  double* pi4l = plam + nproc;        // 1 / (4 lambda)
  double* pvs = pi4l + nproc;         // V_ab + field . hop vector / 2
  double* pvec = pvs + nproc;         // hop vector (3 nproc)
  int32_t* pn = reinterpret_cast<int32_t*>(pvec + 3 * nproc);  // target n
  int32_t* occ = pn + nproc;          // nc
  // compact instance: packed codes of the carriers' sites and of each
  // process's target
  uint32_t* opk = reinterpret_cast<uint32_t*>(occ + nc);  // nc
  uint32_t* pnpk = opk + nc;                              // nproc
  __shared__ double sh_t, sh_energy, sh_dg;
  __shared__ double sh_exp[64];  // 2^(j/64)
  __shared__ int64_t sh_start, sh_hops, sh_row_lo, sh_row_hi, sh_dp_hi;
  __shared__ int32_t sh_stop, sh_noselect, sh_mover;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // the site-dependent part of process p's rate, from the slot tables
  auto load_process = [&](int p) {
    const int c = p / z, j = p - c * z;
    const int s = occ[c];
    const int64_t sj = int64_t(s) * d.max_z + j;
    const int n = d.nbr[sj];
    pn[p] = n;
    if constexpr (kCompact) pnpk[p] = d.pk[n];
    pdphi[p] = d.dphi[sj];
    ptq[p] = d.cq[c] * d.tself[sj];
    pdE[p] = d.E[n] - d.E[s];
    {
      // the same arithmetic as the one-wavefront kernel's slot records
#pragma clang fp contract(off)
      const double lam = d.lam[sj];
      double shift = 0.0;
      if (d.field_on)
        shift = 0.5 * (d.field[0] * d.vec[3 * sj] +
                       d.field[1] * d.vec[3 * sj + 1] +
                       d.field[2] * d.vec[3 * sj + 2]);
      plam[p] = lam;
      pi4l[p] = 1.0 / (4 * lam);
      pvs[p] = d.vab[sj] + shift;
    }
    pvec[3 * p] = d.vec[3 * sj];
    pvec[3 * p + 1] = d.vec[3 * sj + 1];
    pvec[3 * p + 2] = d.vec[3 * sj + 2];
  };
  for (int c = tid; c < nc; c += kThreads) {
    occ[c] = d.occ[c];
    if constexpr (kCompact) opk[c] = d.pk[occ[c]];
  }
  for (int k = tid; k < 3 * nc; k += kThreads) {
    disp[k] = d.disp[k];
    base[k] = d.base[k];
    drift[k] = d.drift[k];
  }
  if (tid < 64) sh_exp[tid] = __builtin_bit_cast(double, kExp2J64[tid]);
  if (tid == 0) {
    sh_t = d.sc->t;
    sh_energy = d.sc->energy;
    sh_start = d.sc->start;
    sh_hops = d.sc->hops;
    sh_noselect = d.sc->no_select;
    sh_stop = d.sc->finished ? 1 : 0;
    sh_mover = -1;
  }
  __syncthreads();
  for (int p = tid; p < nproc; p += kThreads) load_process(p);
  __syncthreads();
  int64_t h = 0;  // hops made in this call
  int stop = sh_stop;
  int code = stop ? 0 : -1;

  while (!stop) {
    if (max_hops > 0 && sh_hops >= max_hops) {
      code = 0;
      break;
    }
    if (h >= n_pairs) {
      code = 1;  // need uniforms
      break;
    }
    // ---- A: psi[x] = sum_c q_c A[x, occ_c] for carrier sites and targets
    for (int r = tid; r < R; r += kThreads) {
      const int c = r / (1 + z);
      const int slot = r - c * (1 + z);
      const int s = occ[c];
      // A[x, y]: the full table's row x, or (compact) the cell-0 row of x's
      // sublattice at y moved by -cell(x); targets from the global slot
      // tables, which the mover's reload below does not touch
      const double* row;
      RowKey key{};
      if constexpr (kCompact) {
        const uint32_t xp =
            slot == 0 ? opk[c] : d.nbr_pk[int64_t(s) * d.max_z + slot - 1];
        key = row_key(xp, d.gp);
        row = d.Ac + int64_t(xp >> 24) * d.N;
      } else {
        const int x = slot == 0 ? s : d.nbr[int64_t(s) * d.max_z + slot - 1];
        row = d.A + int64_t(x) * d.N;
      }
      auto col = [&](int c3) -> uint32_t {
        if constexpr (kCompact) return rel_col(key, opk[c3], d.gp);
        else return uint32_t(occ[c3]);
      };
      // the row's gathers in flight 16 (then 8) at a time, summed in
      // carrier order
      double acc = 0.0;
      int c2 = 0;
      for (; c2 + 16 <= nc; c2 += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = row[col(c2 + u)];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = fma(d.cq[c2 + u], v[u], acc);
      }
      for (; c2 + 8 <= nc; c2 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = row[col(c2 + u)];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fma(d.cq[c2 + u], v[u], acc);
      }
      for (; c2 < nc; ++c2) acc = fma(d.cq[c2], row[col(c2)], acc);
      psi[r] = acc;
    }
    // the last hop's mover: its processes reloaded for its new site, by
    // the last z threads while the gathers above are in flight
    if (sh_mover >= 0 && tid >= kThreads - z)
      load_process(sh_mover * z + (kThreads - 1 - tid));
    __syncthreads();
    // ---- B: rates
    for (int p = tid; p < nproc; p += kThreads) {
      // products and sums rounded one by one, as NumPy forms them
#pragma clang fp contract(off)
      const int c = p / z;
      const int j = p - c * z;
      const double qs = d.cq[c];
      const double term01 =
          pdphi[p] + (psi[c * (1 + z) + 1 + j] - psi[c * (1 + z)]);
      const double term02 = ptq[p];
      const double g0 = 2 * qs * (term01 + term02) + pdE[p];
      kk[p] = marcus_rate(g0, plam[p], pi4l[p], pvs[p], d.nkt, d.nu, sh_exp);
      dg[p] = g0;
    }
    __syncthreads();
    // ---- C: selection, clock and state update on wavefront 0
    if (tid < 64) {
      const int per = (nproc + 63) / 64;
      const int p0 = lane * per;
      const int p1 = min(p0 + per, nproc);
      // the first process whose cumulative rate exceeds u1 k_tot (the
      // reference compares the cumulative sum of k / k_tot with u1: the
      // same choice but where u1 lies within an ulp of a boundary)
      double ksum = 0.0;
      for (int p = p0; p < p1; ++p) ksum += kk[p];
      const double incl = wave_incl_scan(ksum);
      const double k_tot = read_lane(incl, 63);
      const double u1 = U[2 * h];
      const double u2 = U[2 * h + 1];
      const double thr = u1 * k_tot;
      double cum = dpp_f64<kDppWaveShr1, 0xf, true>(incl);
      int found = -1;
      for (int p = p0; p < p1; ++p) {
        cum += kk[p];
        if (cum > thr) {
          found = p;
          break;
        }
      }
      const unsigned long long mask = __ballot(found >= 0);
      int sel;
      if (mask) {
        const int src = __ffsll(static_cast<long long>(mask)) - 1;
        sel = __shfl(found, src, 64);
      } else {
        sel = nproc - 1;
      }
      if (lane == 0) {
#pragma clang fp contract(off)
        if (!mask) ++sh_noselect;
        const int c = sel / z;
        const int s_new = pn[sel];
        const double* v = pvec + 3 * sel;
        double t = sh_t;
        int64_t end;
        if (u2 > 0.0) {
          t -= LU[h] / k_tot;
          const double q = t / d.dt;
          end = q < 9.0e18 ? int64_t(q) : d.P;
        } else {
          t = INFINITY;
          end = d.P;
        }
        sh_t = t;
        const double g = dg[sel];
        sh_dg = g;
        // delg_0 path: [start, end) clipped to P (numpy slice semantics)
        sh_dp_hi = min(end, d.P);
        occ[c] = s_new;
        if constexpr (kCompact) opk[c] = pnpk[sel];
        sh_mover = c;
        disp[3 * c] += v[0];
        disp[3 * c + 1] += v[1];
        disp[3 * c + 2] += v[2];
        if (d.field_on) {
          drift[3 * c] += v[0] * kk[sel];
          drift[3 * c + 1] += v[1] * kk[sel];
          drift[3 * c + 2] += v[2] * kk[sel];
        }
        sh_energy += g;
        if (hop_time) hop_time[h] = t;
        if (hop_proc) hop_proc[h] = sel;
        sh_row_lo = sh_start;
        sh_row_hi = sh_start;
        if (end >= sh_start + 1) {
          if (end >= d.P) end = d.P;
          sh_row_hi = end;
        }
        sh_hops += 1;
        // loop condition of the reference: continue while end < P
        sh_stop = (sh_row_hi > sh_row_lo && sh_row_hi >= d.P) ? 1 : 0;
      }
    }
    __syncthreads();
    ++h;
    // ---- D: path-step outputs on all threads, for a hop that ends one or
    // more path steps (otherwise nothing is written and no barrier is due:
    // the next hop's first two barriers order C's shared values)
    if (sh_row_hi > sh_row_lo) {
      const int64_t lo = sh_row_lo, hi = sh_row_hi;
      const double g = sh_dg;
      for (int64_t r = lo + tid; r < sh_dp_hi; r += kThreads) d.dpath[r] = g;
      if (hi > lo) {
        const int w = 3 * nc;
        const double e = sh_energy;
        for (int64_t idx = tid; idx < (hi - lo) * w; idx += kThreads) {
          const int64_t r = lo + idx / w;
          const int k = int(idx % w);
          d.rows[r * w + k] = base[k] + disp[k];
        }
        for (int64_t r = lo + tid; r < hi; r += kThreads) d.epath[r] = e;
      }
      __syncthreads();
      for (int k = tid; k < 3 * nc; k += kThreads) {
        base[k] += disp[k];
        disp[k] = 0.0;
      }
      if (tid == 0) sh_start = hi;
      __syncthreads();
    }
    stop = sh_stop;
    if (stop) code = 0;
  }
  // ---- write the state back
  for (int c = tid; c < nc; c += kThreads) d.occ[c] = occ[c];
  for (int k = tid; k < 3 * nc; k += kThreads) {
    d.disp[k] = disp[k];
    d.base[k] = base[k];
    d.drift[k] = drift[k];
  }
  if (tid == 0) {
    d.sc->t = sh_t;
    d.sc->energy = sh_energy;
    d.sc->start = sh_start;
    d.sc->hops = sh_hops;
    d.sc->no_select = sh_noselect;
    if (stop) d.sc->finished = 1;
    status[0] = code;
    status[1] = int32_t(h);
  }
}

// ---------------------------------------------------------------------------
// Inputs of the one-wavefront kernel.  Per (site, slot) a 64-byte record:
// the self term A[s, s] - A[s, n], lambda, 1 / (4 lambda), V_ab plus the
// field shift 0.5 F.d, the hop vector d and the lattice part
// q_L . (A[n, :] - A[s, :]) of the hop's potential difference; per site
// (phi_L, E); per (site, slot) the target and its packed code (0 without
// cell-0 rows).
struct alignas(16) SlotRec {
  double tself, lam, i4l, vs;
  double v[3];
  double dphi;
};

// What the one-wavefront kernel's hop loop reads, by value; state read only
// before and after the loop comes from a KmcDev in device memory.
struct WaveArgs {
  const char* tab;       // the table (full, or the rows of cell 0)
  uint32_t row_bytes;    // N sizeof(double)
  int32_t uq;            // every carrier has charge q0
  GeoPk gp;
  const SlotRec* rec;
  const double2* pe;
  const int2* nbr2;
  double* rows;
  double* epath;
  double* dpath;
  const double* U;
  const double* LU;
  double* hop_time;
  int32_t* hop_proc;
  int32_t nc, z, mz, field_on;
  double q0, nu, nkt, dt;  // nkt = -1 / kT
  int64_t P, n_pairs, max_hops;
};

// Records of every (site, slot) from the slot tables (the workgroup
// kernel's arithmetic: products and sums rounded one by one).
__global__ void pack_slot_records(int64_t NM, const double* __restrict__ tself,
                                  const double* __restrict__ dphi,
                                  const double* __restrict__ lam,
                                  const double* __restrict__ vab,
                                  const double* __restrict__ vec,
                                  const int32_t* __restrict__ nbr,
                                  const uint32_t* __restrict__ nbr_pk,
                                  double f0, double f1, double f2,
                                  int field_on, SlotRec* __restrict__ rec,
                                  int2* __restrict__ nbr2) {
#pragma clang fp contract(off)
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= NM) return;
  SlotRec r;
  r.tself = tself[k];
  r.lam = lam[k];
  r.i4l = 1.0 / (4 * r.lam);
  r.v[0] = vec[3 * k];
  r.v[1] = vec[3 * k + 1];
  r.v[2] = vec[3 * k + 2];
  const double shift =
      field_on ? 0.5 * (f0 * r.v[0] + f1 * r.v[1] + f2 * r.v[2]) : 0.0;
  r.vs = vab[k] + shift;
  r.dphi = dphi[k];
  rec[k] = r;
  nbr2[k] = make_int2(nbr[k], nbr_pk ? int(nbr_pk[k]) : 0);
}

__global__ void pack_site_pe(int64_t N, const double* __restrict__ phiL,
                             const double* __restrict__ E,
                             double2* __restrict__ pe) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < N) pe[k] = make_double2(phiL[k], E[k]);
}

// One-wavefront instance with a row cache (used when it fits: n_c <= 64,
// n_c z <= 256, z <= 8, the cache within the LDS of a workgroup and the
// table below 4 GiB).
//
// Rows: process row p (p < n_c z) is the target site of process p, site row
// c is carrier c's site.  The cache holds G[c][row] = A[x_row, occ_c] for
// every carrier c and row, so psi[row] = sum_c q_c G[c][row] is summed from
// LDS in carrier order: the FMA chain the workgroup kernel forms from its
// gathers, so both give the same psi bit for bit.  When carrier m hops
// s -> n only column m of the other rows changes (A[x_row, n], one gather
// per row) and the mover's 1 + z rows are new (n_c gathers each): a hop
// takes (1 + z)(2 n_c - 1) gathers instead of n_c (1 + z) n_c, all issued
// together.  No gather address waits for another load: every process keeps
// the z targets of its own target site (loaded when the process is made),
// so the mover's new rows are in LDS when it is selected.
//
// Lane l owns processes l kNpl .. l kNpl + kNpl - 1 (the workgroup kernel's
// split of the selection, so k_tot and the prefix are summed in the same
// order) with their rate inputs in registers, site row l and carrier l's
// displacement, path base and drift (l < n_c).  One wavefront: no workgroup
// barriers; __syncthreads only waits for the wavefront's own LDS
// operations.  Cheap per-lane work is done by every lane and kept by a
// select rather than branched around; gathers of lanes that have no entry
// to update read a valid address and are dropped.
template <typename T>
using gmem = const __attribute__((address_space(1))) T*;

template <bool kCompact, int kNpl>
__global__ void __launch_bounds__(64)
kmc_wave_kernel(WaveArgs a, const KmcDev* __restrict__ dv,
                int32_t* __restrict__ status) {
  constexpr int kB = 5;  // (1 + z) n_c <= 320 mover-row gathers: 5 per lane
  extern __shared__ double smem[];
  const int nc = a.nc, z = a.z, nproc = nc * z;
  const int R = nproc + nc;
  // carrier pairs of the cache, rounded up to whole groups of 4 (the
  // padding's entries and charges are zero)
  const int ncp = (((nc + 1) >> 1) + 3) & ~3;
  const int lane = threadIdx.x;
  __shared__ double ext[64];  // 2^(j/64)
  ext[lane] = __builtin_bit_cast(double, kExp2J64[lane]);
  // cache: carriers 2p and 2p + 1 of a row side by side, so that one
  // 16-byte read serves two terms of a row's sum
  double2* G = reinterpret_cast<double2*>(smem);          // ncp x R
  double* psis = reinterpret_cast<double*>(G + size_t(ncp) * R);  // nc
  double* cqs = psis + nc;                                // 2 ncp
  int32_t* occ = reinterpret_cast<int32_t*>(cqs + 2 * ncp);  // nc
  uint32_t* opk = reinterpret_cast<uint32_t*>(occ + nc);  // nc
  int32_t* tt = reinterpret_cast<int32_t*>(opk + nc);     // nproc z
  uint32_t* ttpk = reinterpret_cast<uint32_t*>(tt + nproc * z);
  auto gset = [&](int c, int row, double v) {
    double* e = reinterpret_cast<double*>(G + size_t(c >> 1) * R + row);
    e[c & 1] = v;
  };

  // A[x, y]: the full table's row x, or (compact) the cell-0 row of x's
  // sublattice at y moved by -cell(x); 32-bit byte offsets from the table's
  // base (the host keeps the table below 4 GiB for this kernel)
  auto entry = [&](int x, uint32_t xk, int y, uint32_t yk) -> double {
    uint32_t off;
    if constexpr (kCompact)
      off = (xk >> 24) * a.row_bytes + rel_col(xk, yk, a.gp) * 8u;
    else
      off = uint32_t(x) * a.row_bytes + uint32_t(y) * 8u;
    return *(gmem<double>)(a.tab + off);
  };
  // psi of this lane's rows (its processes' target rows and its site row):
  // for each row the FMA chain over carriers 0 .. n_c - 1, the workgroup
  // kernel's chain (the zero padding adds fma(q, 0, acc) = acc).  The cache
  // entries of all the rows are read kC carrier pairs at a time before
  // they are summed, so a group waits for the LDS once.  Rows of lanes
  // without that process or carrier read row 0 and are not kept.
  constexpr int kRows = kNpl + 1;
  constexpr int kC = kRows <= 2 ? 4 : 2;  // divides the padded pair count
  auto row_sums = [&](const int (&row)[kRows], double (&out)[kRows]) {
    double acc[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) acc[r] = 0.0;
    auto sum = [&](auto uniform_q) {
      constexpr bool kUq = decltype(uniform_q)::value;
      for (int p0 = 0; p0 < ncp; p0 += kC) {
        double2 v[kC][kRows];
#pragma unroll
        for (int u = 0; u < kC; ++u)
#pragma unroll
          for (int r = 0; r < kRows; ++r)
            v[u][r] = G[size_t(p0 + u) * R + row[r]];
#pragma unroll
        for (int u = 0; u < kC; ++u) {
          const double q0 = kUq ? a.q0 : cqs[2 * (p0 + u)];
          const double q1 = kUq ? a.q0 : cqs[2 * (p0 + u) + 1];
#pragma unroll
          for (int r = 0; r < kRows; ++r) {
            acc[r] = fma(q0, v[u][r].x, acc[r]);
            acc[r] = fma(q1, v[u][r].y, acc[r]);
          }
        }
      }
    };
    if (a.uq) sum(std::true_type{});
    else sum(std::false_type{});
#pragma unroll
    for (int r = 0; r < kRows; ++r) out[r] = acc[r];
  };

  // ---- processes of this lane: carrier, slot, target, rate inputs
  int pc[kNpl], pj[kNpl], px[kNpl];
  uint32_t pxk[kNpl];
  bool pv[kNpl];
  double pqs[kNpl], pdphi[kNpl], pts[kNpl], pEn[kNpl], pEs[kNpl], plam[kNpl],
      pi4l[kNpl], pvs[kNpl], pvx[kNpl], pvy[kNpl], pvz[kNpl], pst[kNpl];
  // the site-dependent part of process i's rate: slot j of site s, whose
  // target is n.  Kept as loaded (the self term and both site energies,
  // combined in the rate), so that nothing waits for these loads before
  // the next hop's rates need them.
  auto load_inputs = [&](int i, int s, int j, int n) {
    const SlotRec& r = a.rec[int64_t(s) * a.mz + j];
    const double2 ps = a.pe[s], pn = a.pe[n];
    pdphi[i] = r.dphi;
    pts[i] = r.tself;
    pEn[i] = pn.y;
    pEs[i] = ps.y;
    plam[i] = r.lam;
    pi4l[i] = r.i4l;
    pvs[i] = r.vs;
    pvx[i] = r.v[0];
    pvy[i] = r.v[1];
    pvz[i] = r.v[2];
  };

  for (int c = lane; c < nc; c += 64) {
    const int s = dv->occ[c];
    occ[c] = s;
    opk[c] = kCompact ? dv->pk[s] : 0u;
    cqs[c] = dv->cq[c];
  }
  for (int c = nc + lane; c < 2 * ncp; c += 64) cqs[c] = 0.0;
  // carrier `lane`'s displacement since the last path row, path base and
  // drift (n_c <= 64)
  double dsp[3] = {0.0, 0.0, 0.0}, bse[3] = {0.0, 0.0, 0.0},
         drf[3] = {0.0, 0.0, 0.0};
  if (lane < nc) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dsp[k] = dv->disp[3 * lane + k];
      bse[k] = dv->base[3 * lane + k];
      drf[k] = dv->drift[3 * lane + k];
    }
  }
  __syncthreads();
  int last_i = -1;
#pragma unroll
  for (int i = 0; i < kNpl; ++i) {
    const int p = lane * kNpl + i;
    pv[i] = p < nproc;
    pc[i] = pv[i] ? p / z : 0;
    pj[i] = pv[i] ? p - pc[i] * z : 0;
    pqs[i] = cqs[pc[i]];
    px[i] = 0;
    pxk[i] = 0u;
    pst[i] = 0.0;
    if (pv[i]) {
      last_i = i;
      const int s = occ[pc[i]];
      const int2 t = a.nbr2[int64_t(s) * a.mz + pj[i]];
      px[i] = t.x;
      pxk[i] = uint32_t(t.y);
      load_inputs(i, s, pj[i], t.x);
    } else {
      pdphi[i] = pts[i] = pEn[i] = pEs[i] = pvs[i] = 0.0;
      plam[i] = 1.0;
      pi4l[i] = 0.25;
      pvx[i] = pvy[i] = pvz[i] = 0.0;
    }
  }
  // every process's look-ahead: the z targets of its target
  for (int u = lane; u < nproc * z; u += 64) {
    const int p = u / z, k = u - p * z;
    const int c = p / z, j = p - c * z;
    const int n = a.nbr2[int64_t(occ[c]) * a.mz + j].x;
    const int2 t = a.nbr2[int64_t(n) * a.mz + k];
    tt[u] = t.x;
    ttpk[u] = uint32_t(t.y);
  }
  // the cache, every (carrier, row); the pairs past n_c stay zero
  for (int u = lane; u < ncp * R; u += 64)
    G[u] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int u = lane; u < nc * R; u += 64) {
    const int row = u / nc, c = u - row * nc;
    int x;
    uint32_t xk;
    if (row < nproc) {
      const int c2 = row / z, j2 = row - c2 * z;
      const int2 t = a.nbr2[int64_t(occ[c2]) * a.mz + j2];
      x = t.x;
      xk = uint32_t(t.y);
    } else {
      x = occ[row - nproc];
      xk = opk[row - nproc];
    }
    gset(c, row, entry(x, xk, occ[c], opk[c]));
  }
  // site row of this lane (lanes past n_c keep carrier 0's, never stored)
  int sx = occ[lane < nc ? lane : 0];
  uint32_t sxk = opk[lane < nc ? lane : 0];
  __syncthreads();
  // this lane's rows and where their sums go
  auto lane_sums = [&]() {
    int row[kRows];
    double out[kRows];
#pragma unroll
    for (int i = 0; i < kNpl; ++i) row[i] = pv[i] ? lane * kNpl + i : 0;
    row[kNpl] = lane < nc ? nproc + lane : 0;
    row_sums(row, out);
#pragma unroll
    for (int i = 0; i < kNpl; ++i)
      if (pv[i]) pst[i] = out[i];
    if (lane < nc) psis[lane] = out[kNpl];
  };
  lane_sums();
  // the mover-row gathers this lane takes: row brl (0: the mover's site,
  // 1 + j: its target j) at carrier bc; nb of the kB rounds hold any
  int brl[kB], bc[kB];
#pragma unroll
  for (int k = 0; k < kB; ++k) {
    const int u = lane + 64 * k;
    brl[k] = u < (1 + z) * nc ? u / nc : -1;
    bc[k] = brl[k] >= 0 ? u - brl[k] * nc : 0;
  }
  const int nb = ((1 + z) * nc + 63) / 64;
  // look-ahead load of this lane: target lk of the mover's new target lj
  const bool la_on = lane < z * z;
  const int lj = la_on ? lane / z : 0;
  const int lk = la_on ? lane - lj * z : 0;
  __syncthreads();

  const KmcScalars sc0 = *dv->sc;
  double t = sc0.t, energy = sc0.energy;
  int64_t start = sc0.start, hops = sc0.hops;
  int noselect = sc0.no_select;
  int stop = sc0.finished ? 1 : 0;
  int code = stop ? 0 : -1;
  int64_t h = 0;
  double u1c = 0.0, u2c = 0.0, luc = 0.0;
  // hop log of the current block of 64 hops: lane l holds hop (h & ~63) + l
  double tlog = 0.0;
  int32_t plog = 0;
  auto flush_log = [&](int64_t block, int count) {
    if (lane < count) {
      if (a.hop_time) a.hop_time[block + lane] = tlog;
      if (a.hop_proc) a.hop_proc[block + lane] = plog;
    }
  };
  while (!stop) {
    if (a.max_hops > 0 && hops >= a.max_hops) {
      code = 0;
      break;
    }
    if (h >= a.n_pairs) {
      code = 1;  // need uniforms
      break;
    }
    const int hl = int(h & 63);
    if (hl == 0) {
      // lane l keeps the uniforms of hop h + l
      const int64_t q = h + lane;
      if (q < a.n_pairs) {
        u1c = a.U[2 * q];
        u2c = a.U[2 * q + 1];
        luc = a.LU[q];
      }
    }
    // ---- rates of this lane's processes
    double kk[kNpl], gg[kNpl];
    double ksum = 0.0;
#pragma unroll
    for (int i = 0; i < kNpl; ++i) {
      // products and sums rounded one by one, as NumPy forms them (and as
      // the workgroup kernel does, whose rates pass through LDS)
#pragma clang fp contract(off)
      kk[i] = 0.0;
      gg[i] = 0.0;
      if (pv[i]) {
        const double term01 = pdphi[i] + (pst[i] - psis[pc[i]]);
        const double g0 =
            2 * pqs[i] * (term01 + pqs[i] * pts[i]) + (pEn[i] - pEs[i]);
        kk[i] = marcus_rate(g0, plam[i], pi4l[i], pvs[i], a.nkt, a.nu, ext);
        gg[i] = g0;
        ksum += kk[i];
      }
    }
    // ---- k_tot and the selection: the first process whose cumulative
    // rate exceeds u1 k_tot (the workgroup kernel's order of sums)
    const double incl = wave_incl_scan(ksum);
    const double k_tot = read_lane(incl, 63);
    const double u1 = read_lane(u1c, hl);
    const double u2 = read_lane(u2c, hl);
    const double lu = read_lane(luc, hl);
    const double thr = u1 * k_tot;
    double cum = dpp_f64<kDppWaveShr1, 0xf, true>(incl);
    int found = -1;
#pragma unroll
    for (int i = 0; i < kNpl; ++i) {
      if (pv[i] && found < 0) {
        cum += kk[i];
        if (cum > thr) found = i;
      }
    }
    const unsigned long long mask = __ballot(found >= 0);
    int src;
    if (mask) {
      src = __ffsll(static_cast<long long>(mask)) - 1;
    } else {
      src = (nproc - 1) / kNpl;  // rounding left no cum > u1: the last one
      ++noselect;
    }
    const int mi = found >= 0 ? found : last_i;
    double k_my = kk[0], g_my = gg[0], vx_my = pvx[0], vy_my = pvy[0],
           vz_my = pvz[0];
    int c_my = pc[0], x_my = px[0];
    uint32_t xk_my = pxk[0];
#pragma unroll
    for (int i = 1; i < kNpl; ++i) {
      if (mi == i) {
        k_my = kk[i];
        g_my = gg[i];
        vx_my = pvx[i];
        vy_my = pvy[i];
        vz_my = pvz[i];
        c_my = pc[i];
        x_my = px[i];
        xk_my = pxk[i];
      }
    }
    const int sel = __builtin_amdgcn_readlane(lane * kNpl + mi, src);
    const int m = __builtin_amdgcn_readlane(c_my, src);
    const int n = __builtin_amdgcn_readlane(x_my, src);
    const uint32_t npk = uint32_t(__builtin_amdgcn_readlane(int(xk_my), src));
    const double k_sel = read_lane(k_my, src);
    const double g_sel = read_lane(g_my, src);
    const double v_sel[3] = {read_lane(vx_my, src), read_lane(vy_my, src),
                             read_lane(vz_my, src)};

    // ---- one round of loads: column m of the other rows, the mover's new
    // rows, the rate inputs of its new processes and their look-ahead
    const int tsel = sel * z;  // the mover's new targets: tt[tsel + j]
    double ga[kNpl];
#pragma unroll
    for (int i = 0; i < kNpl; ++i) ga[i] = entry(px[i], pxk[i], n, npk);
    const double gs_site = entry(sx, sxk, n, npk);
    // every round's LDS inputs read first, then every round's gather
    // issued, unused rounds included (they read a valid entry that is not
    // kept): no load then sits behind another round's LDS wait, and every
    // load lands in its own register, so nothing waits for the gathers
    // before the cache is written
    int gx[kB], gy[kB];
    uint32_t gxk[kB], gyk[kB];
#pragma unroll
    for (int k = 0; k < kB; ++k) {
      const int rl = brl[k] > 0 ? brl[k] : 0;
      const int ti = tsel + (rl > 0 ? rl - 1 : 0);
      const int tx = tt[ti];
      const uint32_t txk = ttpk[ti];
      const int cc = bc[k];
      const int oy = occ[cc];
      const uint32_t oyk = opk[cc];
      gx[k] = rl > 0 ? tx : n;
      gxk[k] = rl > 0 ? txk : npk;
      gy[k] = cc == m ? n : oy;
      gyk[k] = cc == m ? npk : oyk;
    }
    double gb[kB];
#pragma unroll
    for (int k = 0; k < kB; ++k) gb[k] = entry(gx[k], gxk[k], gy[k], gyk[k]);
    const int2 la = a.nbr2[int64_t(tt[tsel + lj]) * a.mz + lk];

    // ---- clock, energy and path bookkeeping (uniform), while the loads
    // are in flight
    int64_t end;
    if (u2 > 0.0) {
      t -= lu / k_tot;
      const double qd = t / a.dt;
      end = qd < 9.0e18 ? int64_t(qd) : a.P;
    } else {
      t = INFINITY;
      end = a.P;
    }
    energy += g_sel;
    tlog = lane == hl ? t : tlog;
    plog = lane == hl ? sel : plog;
    const int64_t row_lo = start;
    int64_t row_hi = start;
    if (end >= start + 1) row_hi = end >= a.P ? a.P : end;
    const int64_t dp_hi = min(end, a.P);
    ++hops;
    // loop condition of the reference: continue while end < P
    stop = (row_hi > row_lo && row_hi >= a.P) ? 1 : 0;
    {
      // the mover's displacement and drift, in lane m
#pragma clang fp contract(off)
      const bool me = lane == m;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        dsp[k] = me ? dsp[k] + v_sel[k] : dsp[k];
        if (a.field_on) drf[k] = me ? drf[k] + v_sel[k] * k_sel : drf[k];
      }
    }
    if (lane == 0) {
      occ[m] = n;
      opk[m] = npk;
    }
    if (lane == m) {
      sx = n;
      sxk = npk;
    }
    // the rate inputs of the mover's new processes (after the uniform work
    // above, so that it overlaps the loads)
#pragma unroll
    for (int i = 0; i < kNpl; ++i) {
      if (pv[i] && pc[i] == m) {
        const int nn = tt[tsel + pj[i]];
        px[i] = nn;
        pxk[i] = ttpk[tsel + pj[i]];
        load_inputs(i, n, pj[i], nn);
      }
    }
    // ---- the cache: column m of the other rows first, then the mover's
    // rows (which overwrite what the first writes put in them)
#pragma unroll
    for (int i = 0; i < kNpl; ++i)
      if (pv[i]) gset(m, lane * kNpl + i, ga[i]);
    if (lane < nc) gset(m, nproc + lane, gs_site);
#pragma unroll
    for (int k = 0; k < kB; ++k) {
      if (k < nb && brl[k] >= 0) {
        const int row = brl[k] == 0 ? nproc + m : m * z + brl[k] - 1;
        gset(bc[k], row, gb[k]);
      }
    }
    if (la_on) {
      tt[(m * z + lj) * z + lk] = la.x;
      ttpk[(m * z + lj) * z + lk] = uint32_t(la.y);
    }
    __syncthreads();
    lane_sums();
    __syncthreads();
    // ---- path-step outputs, for a hop that ends one or more path steps
    if (row_hi > row_lo) {
      for (int64_t r = row_lo + lane; r < dp_hi; r += 64) a.dpath[r] = g_sel;
      for (int64_t r = row_lo + lane; r < row_hi; r += 64) a.epath[r] = energy;
      if (lane < nc) {
        const int w = 3 * nc;
        for (int64_t r = row_lo; r < row_hi; ++r)
#pragma unroll
          for (int k = 0; k < 3; ++k) a.rows[r * w + 3 * lane + k] =
              bse[k] + dsp[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          bse[k] += dsp[k];
          dsp[k] = 0.0;
        }
      }
      start = row_hi;
    }
    if (hl == 63) flush_log(h - 63, 64);
    ++h;
    if (stop) code = 0;
  }
  // ---- write the state back
  flush_log(h & ~int64_t(63), int(h & 63));
  __syncthreads();
  for (int c = lane; c < nc; c += 64) dv->occ[c] = occ[c];
  if (lane < nc) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dv->disp[3 * lane + k] = dsp[k];
      dv->base[3 * lane + k] = bse[k];
      dv->drift[3 * lane + k] = drf[k];
    }
  }
  if (lane == 0) {
    KmcScalars o = sc0;
    o.t = t;
    o.energy = energy;
    o.start = start;
    o.hops = hops;
    o.no_select = noselect;
    if (stop) o.finished = 1;
    *dv->sc = o;
    status[0] = code;
    status[1] = int32_t(h);
  }
}

// Scratch of pycd_kmc_advance_static: log(u2) of up to kStaticPairs uniform
// pairs, then the hop loop's two status words.  One copy per device (every
// device the library runs on holds it until the process ends); host code
// reaches it only through the address hipGetSymbolAddress returns for the
// device (pycd_ctx_create resolves it once per device).  Calls that overlap
// on one device share it and race.
constexpr int64_t kStaticPairs = int64_t(1) << 20;
__device__ __align__(256) double scratch[kStaticPairs + 32];

__global__ void reset_status(int32_t* __restrict__ status) {
  status[0] = -1;
  status[1] = 0;
}

// phi_L[i] = sum_j A[i, j] q_L[j]: one wavefront per row.
__global__ void __launch_bounds__(256)
row_dot_kernel(const double* __restrict__ A, const double* __restrict__ q,
               int64_t N, double* __restrict__ out) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t i = int64_t(blockIdx.x) * 4 + wave; i < N;
       i += int64_t(gridDim.x) * 4) {
    const double* row = A + i * N;
    double acc = 0.0;
    for (int64_t j = lane; j < N; j += 64) acc = fma(row[j], q[j], acc);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) out[i] = acc;
  }
}

// tself[s, j] = A[s, s] - A[s, nbr(s, j)] for valid slots.
__global__ void self_term_kernel(const double* __restrict__ A,
                                 const int32_t* __restrict__ nbr, int64_t N,
                                 int max_z, double* __restrict__ tself) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= N * max_z) return;
  const int64_t s = idx / max_z;
  const int n = nbr[idx];
  tself[idx] = n >= 0 ? A[s * N + s] - A[s * N + n] : 0.0;
}

// Lattice part of the potential difference of each hop, the reference's
// q . (A[n, :] - A[s, :]) restricted to the lattice charges: every product
// q_L[y] (A[n, y] - A[s, y]) rounded as NumPy forms it, their sum kept as
// an unevaluated pair (hi, lo) (Knuth's TwoSum) and rounded once at the
// end.  Forming it as phi_L[n] - phi_L[s] instead would carry about an ulp
// of phi_L, a sum of order 10 Ha, into every rate.  One wavefront per
// (site, slot); 0 for empty slots.
__device__ inline void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__global__ void __launch_bounds__(256)
slot_dphi_kernel(const double* __restrict__ A, const double* __restrict__ q,
                 const int32_t* __restrict__ nbr, int64_t N, int max_z,
                 double* __restrict__ dphi) {
#pragma clang fp contract(off)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t NM = N * max_z;
  for (int64_t k = int64_t(blockIdx.x) * 4 + wave; k < NM;
       k += int64_t(gridDim.x) * 4) {
    const int n = nbr[k];
    if (n < 0) {
      if (lane == 0) dphi[k] = 0.0;
      continue;
    }
    const double* rs = A + (k / max_z) * N;
    const double* rn = A + int64_t(n) * N;
    double hi = 0.0, lo = 0.0;
    for (int64_t y = lane; y < N; y += 64) {
      const double p = q[y] * (rn[y] - rs[y]);
      double e;
      two_sum(hi, p, hi, e);
      lo += e;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double h2 = __shfl_xor(hi, off, 64);
      const double l2 = __shfl_xor(lo, off, 64);
      double e;
      two_sum(hi, h2, hi, e);
      lo = lo + l2 + e;
    }
    if (lane == 0) dphi[k] = hi + lo;
  }
}

// Initial energy pieces: e[0] = q_L . phi_L is precomputed on the host side
// of the session; this kernel adds the carrier terms
// 2 sum_c q_c phi_L[occ_c] + sum_{c,c'} q_c q_c' A[occ_c, occ_c'].
__global__ void carrier_energy_kernel(const double* __restrict__ A,
                                      const double* __restrict__ phiL,
                                      const double* __restrict__ cq,
                                      const int32_t* __restrict__ occ,
                                      int nc, int64_t N,
                                      double* __restrict__ out) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int c = threadIdx.x; c < nc; c += blockDim.x) {
    const double* row = A + int64_t(occ[c]) * N;
    double inner = 0.0;
    for (int c2 = 0; c2 < nc; ++c2) inner = fma(cq[c2], row[occ[c2]], inner);
    acc += cq[c] * (2.0 * phiL[occ[c]] + inner);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

__global__ void dot_kernel(const double* __restrict__ a,
                           const double* __restrict__ b, int64_t n,
                           double* __restrict__ out) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x)
    acc = fma(a[k], b[k], acc);
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

}  // namespace pycd

using namespace pycd;

struct pycd_kmc {
  pycd_ctx* ctx = nullptr;
  pycd_kmc_problem pb{};
  DevBuf<double> A, phiL, dphi, vec, lam, vab, tself, E, cq, qL;
  DevBuf<int32_t> nbr, occ, status;
  DevBuf<double> disp, base, drift, rows, epath, dpath, tmp;
  DevBuf<KmcScalars> sc;
  DevBuf<double> ubuf, tbuf, lubuf;
  DevBuf<int32_t> pbuf;
  // periodic supercell: cell-0 rows of the table and packed site codes
  bool compact = false;
  Geo geo{};
  DevBuf<double> Ac;
  DevBuf<uint32_t> pk, nbr_pk;
  // one-wavefront kernel: slot records, (phi_L, E) per site, targets with
  // their codes, the device copy of the kernel's state pointers
  DevBuf<SlotRec> rec;
  DevBuf<double2> pe;
  DevBuf<int2> nbr2;
  DevBuf<KmcDev> dev;
  double qLAqL = 0.0;
  size_t smem = 0;
  // one-wavefront row-cache kernel: whether it fits, its LDS, processes
  // per lane, and which kernel the hop loop runs (PYCD_KMC_*)
  bool wave_ok = false;
  size_t wave_smem = 0;
  int wave_npl = 0;
  int32_t variant = PYCD_KMC_AUTO;
  bool uniform_q = false;  // every carrier has charge q0
  double q0 = 0.0;
  bool begun = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  ~pycd_kmc() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

namespace {

int check_problem(pycd_ctx* ctx, const pycd_kmc_problem* pb) {
  if (!pb) return fail(ctx, PYCD_ERR_ARG, "null problem");
  if (pb->n_sites <= 0 || pb->n_sites > (int64_t(1) << 31) - 1)
    return fail(ctx, PYCD_ERR_ARG, "bad n_sites");
  if (pb->n_carriers <= 0 || pb->z <= 0 || pb->z > 64 || pb->max_z < pb->z)
    return fail(ctx, PYCD_ERR_ARG, "bad carrier/slot counts (z <= 64)");
  if (!pb->table || !pb->lattice_q || !pb->nbr_site || !pb->nbr_vec ||
      !pb->nbr_lambda || !pb->nbr_vab || !pb->site_energy || !pb->carrier_q)
    return fail(ctx, PYCD_ERR_ARG, "null table pointer");
  if (!(pb->time_interval > 0) || pb->n_path < 1 || !(pb->kT > 0))
    return fail(ctx, PYCD_ERR_ARG, "bad time grid or temperature");
  return PYCD_OK;
}

// phi_lattice = A q_lattice, the lattice part of every hop's potential
// difference and q_lattice^T A q_lattice from the resident table and
// lattice charges.
hipError_t lattice_terms(pycd_kmc* s) {
  hipStream_t st = s->ctx->stream;
  const int64_t N = s->pb.n_sites;
  const int blocks = int(std::min<int64_t>((N + 3) / 4, 8192));
  row_dot_kernel<<<blocks, 256, 0, st>>>(s->A.ptr, s->qL.ptr, N, s->phiL.ptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t NM = N * s->pb.max_z;
  slot_dphi_kernel<<<int(std::min<int64_t>((NM + 3) / 4, 16384)), 256, 0,
                     st>>>(s->A.ptr, s->qL.ptr, s->nbr.ptr, N, s->pb.max_z,
                           s->dphi.ptr);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  dot_kernel<<<1, 256, 0, st>>>(s->qL.ptr, s->phiL.ptr, N, s->tmp.ptr);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(&s->qLAqL, s->tmp.ptr, sizeof(double),
                     hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

// Largest |A[i, j] - Ac[sub(i), rel(i, j)]| over the full table, as the bit
// pattern of a non-negative double (ordered like the values).  One
// workgroup of 256 per 4 rows.
__global__ void __launch_bounds__(256)
periodic_gap(const double* __restrict__ A, const double* __restrict__ Ac,
             int64_t N, Geo g, unsigned long long* __restrict__ worst) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const GeoPk gp = geo_pk(g);
  double m = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * 4 + wave; i < N;
       i += int64_t(gridDim.x) * 4) {
    const uint32_t a = site_pack(uint32_t(i), g);
    const RowKey key = row_key(a, gp);
    const double* row = A + i * N;
    const double* crow = Ac + int64_t(a >> 24) * N;
    for (int64_t j = lane; j < N; j += 64)
      m = nan_max(m, fabs(row[j] - crow[rel_col(key, site_pack(uint32_t(j), g),
                                                 gp)]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    m = nan_max(m, __shfl_xor(m, off, 64));
  if (lane == 0)
    atomicMax(worst, static_cast<unsigned long long>(__double_as_longlong(m)));
}

// A periodic supercell's table is translation-invariant: the hop loop can
// read the n_unit rows of cell 0 (L2-resident) instead of the full table.
// Used only when the cell shape is given, fits the packed site codes, the
// full table is above kCellRowsFrom bytes, and every entry of the full
// table equals its cell-0 translate to kTranslationTol of the table's
// scale (a NaN anywhere fails);
// otherwise the loop reads the full table.  Smaller tables stay in the
// caches whole, and there the column arithmetic costs more than it saves
// (768 sites: 3.51 us per hop on the full table, 3.65 on cell-0 rows;
// 6144 sites, 302 MB: 3.94 and 3.60).
constexpr int64_t kCellRowsFrom = int64_t(32) << 20;
hipError_t compact_setup(pycd_kmc* s, const std::vector<int32_t>& nb) {
  const pycd_kmc_problem& pb = s->pb;
  const int64_t N = pb.n_sites;
  const int mz = pb.max_z;
  const Geo g{pb.cells[0], pb.cells[1], pb.cells[2], pb.n_unit};
  s->compact = false;
  if (g.nx < 1 || g.ny < 1 || g.nz < 1 || g.nu < 1 || g.nx > 255 ||
      g.ny > 255 || g.nz > 255 || g.nu > 255 ||
      int64_t(g.nx) * g.ny * g.nz * g.nu != N || N >= (int64_t(1) << 24) ||
      int64_t(g.ny) * g.nz * g.nu >= 65536 ||
      N * N * int64_t(sizeof(double)) <= kCellRowsFrom)
    return hipSuccess;
  hipStream_t st = s->ctx->stream;
  std::vector<uint32_t> pk(N), npk(size_t(N) * mz, 0u);
  for (int64_t i = 0; i < N; ++i) pk[i] = site_pack(uint32_t(i), g);
  for (size_t k = 0; k < npk.size(); ++k)
    if (nb[k] >= 0) npk[k] = pk[nb[k]];
  const size_t rows = size_t(g.nu) * size_t(N);
  hipError_t e;
  if ((e = s->pk.alloc(N)) != hipSuccess) return e;
  if ((e = s->nbr_pk.alloc(npk.size())) != hipSuccess) return e;
  if ((e = s->Ac.alloc(rows)) != hipSuccess) return e;
  if ((e = upload(s->pk.ptr, pk.data(), pk.size(), st)) != hipSuccess) return e;
  if ((e = upload(s->nbr_pk.ptr, npk.data(), npk.size(), st)) != hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(s->Ac.ptr, s->A.ptr, rows * sizeof(double),
                          hipMemcpyDeviceToDevice, st)) != hipSuccess)
    return e;
  DevBuf<unsigned long long> worst;
  if ((e = worst.alloc(1)) != hipSuccess) return e;
  const unsigned long long zero = 0;
  if ((e = upload(worst.ptr, &zero, 1, st)) != hipSuccess) return e;
  periodic_gap<<<int(std::min<int64_t>((N + 3) / 4, 8192)), 256, 0, st>>>(
      s->A.ptr, s->Ac.ptr, N, g, worst.ptr);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  unsigned long long bits = 0;
  std::vector<double> c0(rows);
  if ((e = hipMemcpyAsync(&bits, worst.ptr, 8, hipMemcpyDeviceToHost, st)) !=
      hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(c0.data(), s->Ac.ptr, rows * sizeof(double),
                          hipMemcpyDeviceToHost, st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  double gap, scale = 0.0;
  std::memcpy(&gap, &bits, 8);
  for (double v : c0) scale = std::max(scale, std::fabs(v));
  if (gap <= kTranslationTol * scale) {
    s->compact = true;
    s->geo = g;
  } else {
    s->Ac.release();
    s->pk.release();
    s->nbr_pk.release();
  }
  return hipSuccess;
}

// Slot records, targets and (phi_L, E) pairs of every site for the
// one-wavefront kernel, from the slot tables and the lattice terms.
hipError_t pack_wave_inputs(pycd_kmc* s) {
  hipStream_t st = s->ctx->stream;
  const int64_t N = s->pb.n_sites;
  const int64_t NM = N * s->pb.max_z;
  const double* f = s->pb.field;
  const bool field_on = f[0] != 0.0 || f[1] != 0.0 || f[2] != 0.0;
  pack_slot_records<<<int((NM + 255) / 256), 256, 0, st>>>(
      NM, s->tself.ptr, s->dphi.ptr, s->lam.ptr, s->vab.ptr, s->vec.ptr,
      s->nbr.ptr, s->compact ? s->nbr_pk.ptr : nullptr, f[0], f[1], f[2],
      field_on ? 1 : 0, s->rec.ptr, s->nbr2.ptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  pack_site_pe<<<int((N + 255) / 256), 256, 0, st>>>(N, s->phiL.ptr,
                                                     s->E.ptr, s->pe.ptr);
  return hipGetLastError();
}

// LDS of the one-wavefront kernel: the row cache (carrier pairs, rounded
// up to groups of 4, x (n_c z + n_c) rows of 16 bytes), site-row psi
// (n_c doubles) and charges (one per padded carrier), carrier sites and
// their codes, and each process's look-ahead (2 n_c z z words).  0 when
// the instance does not apply.
size_t wave_lds_bytes(int nc, int z) {
  const int nproc = nc * z;
  if (nc > 64 || nproc > 256 || z > 8) return 0;
  const size_t ncp = (((size_t(nc) + 1) / 2) + 3) & ~size_t(3);
  return 16 * ncp * (nproc + nc) + sizeof(double) * (size_t(nc) + 2 * ncp) +
         sizeof(int32_t) * (2 * size_t(nc) + 2 * size_t(nproc) * z);
}

}  // namespace

extern "C" int pycd_kmc_create(pycd_ctx* ctx, const pycd_kmc_problem* pb,
                               pycd_kmc** out) {
  if (!ctx || !out) return PYCD_ERR_ARG;
  *out = nullptr;
  int rc = check_problem(ctx, pb);
  if (rc) return rc;
  DeviceGuard guard(ctx->device);
  const int64_t N = pb->n_sites;
  const int nc = pb->n_carriers, z = pb->z, mz = pb->max_z;
  const size_t smem = sizeof(double) * (size_t(nc) * (1 + z) + 11 * size_t(nc) * z +
                                        9 * size_t(nc)) +
                      sizeof(int32_t) * (size_t(nc) * z + nc) +
                      sizeof(uint32_t) * (size_t(nc) * z + nc) + 16;
  if (smem > 60 * 1024)
    return fail(ctx, PYCD_ERR_ARG,
                "n_c=%d, z=%d need %zu B of LDS (limit 60 KiB)", nc, z, smem);
  // Host-side validation of the slot table (every index the kernel reads).
  std::vector<int32_t> nb(size_t(N) * mz);
  {
    if (is_device_ptr(pb->nbr_site)) {
      PYCD_HIP(ctx, hipMemcpy(nb.data(), pb->nbr_site, nb.size() * 4,
                              hipMemcpyDeviceToHost));
    } else {
      std::memcpy(nb.data(), pb->nbr_site, nb.size() * 4);
    }
    for (size_t k = 0; k < nb.size(); ++k)
      if (nb[k] < -1 || nb[k] >= N)
        return fail(ctx, PYCD_ERR_ARG, "nbr_site[%zu]=%d out of range", k,
                    nb[k]);
    // Closure: a hopping site (slot 0 valid) has z valid slots and every
    // target is itself a hopping site, so a carrier never reaches a site
    // whose slot rows the kernel cannot read.
    for (int64_t s = 0; s < N; ++s) {
      const int32_t* row = nb.data() + size_t(s) * mz;
      if (row[0] < 0) continue;
      for (int j = 0; j < z; ++j) {
        if (row[j] < 0)
          return fail(ctx, PYCD_ERR_ARG, "site %lld has fewer than %d slots",
                      (long long)s, z);
        if (nb[size_t(row[j]) * mz] < 0)
          return fail(ctx, PYCD_ERR_ARG,
                      "slot %d of site %lld targets non-hopping site %d", j,
                      (long long)s, row[j]);
      }
    }
  }
  pycd_kmc* s = new (std::nothrow) pycd_kmc();
  if (!s) return PYCD_ERR_NOMEM;
  s->ctx = ctx;
  s->pb = *pb;
  s->smem = smem;
  hipStream_t st = ctx->stream;
  auto bail = [&](int code) {
    delete s;
    return code;
  };
#define S_HIP(expr)                                                         \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return bail(fail(ctx, PYCD_ERR_HIP, "%s: %s", #expr,                  \
                       hipGetErrorString(e_)));                             \
  } while (0)
  const size_t NM = size_t(N) * mz;
  const int64_t P = pb->n_path;
  S_HIP(s->A.alloc(size_t(N) * N));
  S_HIP(s->phiL.alloc(N));
  S_HIP(s->dphi.alloc(NM));
  S_HIP(s->qL.alloc(N));
  S_HIP(s->nbr.alloc(NM));
  S_HIP(s->vec.alloc(NM * 3));
  S_HIP(s->lam.alloc(NM));
  S_HIP(s->vab.alloc(NM));
  S_HIP(s->tself.alloc(NM));
  S_HIP(s->E.alloc(N));
  S_HIP(s->cq.alloc(nc));
  S_HIP(s->occ.alloc(nc));
  S_HIP(s->status.alloc(2));
  S_HIP(s->disp.alloc(3 * nc));
  S_HIP(s->base.alloc(3 * nc));
  S_HIP(s->drift.alloc(3 * nc));
  S_HIP(s->rows.alloc(size_t(P) * 3 * nc));
  S_HIP(s->epath.alloc(P));
  S_HIP(s->dpath.alloc(P));
  S_HIP(s->tmp.alloc(2));
  S_HIP(s->sc.alloc(1));
  S_HIP(upload(s->A.ptr, pb->table, size_t(N) * N, st));
  S_HIP(upload(s->qL.ptr, pb->lattice_q, N, st));
  S_HIP(upload(s->nbr.ptr, pb->nbr_site, NM, st));
  S_HIP(upload(s->vec.ptr, pb->nbr_vec, NM * 3, st));
  S_HIP(upload(s->lam.ptr, pb->nbr_lambda, NM, st));
  S_HIP(upload(s->vab.ptr, pb->nbr_vab, NM, st));
  S_HIP(upload(s->E.ptr, pb->site_energy, N, st));
  S_HIP(upload(s->cq.ptr, pb->carrier_q, nc, st));
  {
    std::vector<double> cq(nc);
    if (is_device_ptr(pb->carrier_q)) {
      S_HIP(hipMemcpy(cq.data(), pb->carrier_q, 8 * size_t(nc),
                      hipMemcpyDeviceToHost));
    } else {
      std::memcpy(cq.data(), pb->carrier_q, 8 * size_t(nc));
    }
    s->q0 = cq[0];
    s->uniform_q = true;
    for (double v : cq) s->uniform_q = s->uniform_q && v == cq[0];
  }
  self_term_kernel<<<int((NM + 255) / 256), 256, 0, st>>>(s->A.ptr, s->nbr.ptr,
                                                         N, mz, s->tself.ptr);
  S_HIP(hipGetLastError());
  S_HIP(lattice_terms(s));
  S_HIP(compact_setup(s, nb));
  S_HIP(s->rec.alloc(NM));
  S_HIP(s->nbr2.alloc(NM));
  S_HIP(s->pe.alloc(size_t(N)));
  S_HIP(s->dev.alloc(1));
  S_HIP(pack_wave_inputs(s));
  {
    int lds_max = 0;
    S_HIP(hipDeviceGetAttribute(&lds_max,
                                hipDeviceAttributeMaxSharedMemoryPerBlock,
                                ctx->device));
    s->wave_smem = wave_lds_bytes(nc, z);
    s->wave_npl = (nc * z + 63) / 64;
    // the kernel addresses the table with 32-bit byte offsets
    const double tab_bytes = 8.0 * double(s->compact ? s->geo.nu : N) * N;
    // (the kernel's own 512 bytes of 2^(j/64) beside the dynamic part)
    s->wave_ok = s->wave_smem > 0 && s->wave_smem + 512 <= size_t(lds_max) &&
                 tab_bytes < 4294967296.0;
  }
#undef S_HIP
  *out = s;
  return PYCD_OK;
}

extern "C" int pycd_kmc_set_variant(pycd_kmc* s, int32_t variant) {
  if (!s) return PYCD_ERR_ARG;
  if (variant != PYCD_KMC_AUTO && variant != PYCD_KMC_WORKGROUP &&
      variant != PYCD_KMC_WAVE)
    return fail(s->ctx, PYCD_ERR_ARG, "unknown kernel variant %d", variant);
  if (variant == PYCD_KMC_WAVE && !s->wave_ok)
    return fail(s->ctx, PYCD_ERR_ARG,
                "the one-wavefront kernel needs n_c <= 64, n_c z <= 256, "
                "z <= 8 and %zu B of LDS (n_c=%d, z=%d)",
                s->wave_smem, s->pb.n_carriers, s->pb.z);
  s->variant = variant;
  return PYCD_OK;
}

extern "C" int pycd_kmc_variant(const pycd_kmc* s, int32_t* variant) {
  if (!s || !variant) return PYCD_ERR_ARG;
  const bool wave = s->variant == PYCD_KMC_WAVE ||
                    (s->variant == PYCD_KMC_AUTO && s->wave_ok);
  *variant = wave ? PYCD_KMC_WAVE : PYCD_KMC_WORKGROUP;
  return PYCD_OK;
}

extern "C" int pycd_kmc_table_rows(const pycd_kmc* s, int64_t* rows) {
  if (!s || !rows) return PYCD_ERR_ARG;
  *rows = s->compact ? int64_t(s->geo.nu) : s->pb.n_sites;
  return PYCD_OK;
}

extern "C" int pycd_kmc_set_sites(pycd_kmc* s, const double* lattice_q,
                                  const double* site_energy) {
  if (!s) return PYCD_ERR_ARG;
  pycd_ctx* ctx = s->ctx;
  if (!lattice_q || !site_energy)
    return fail(ctx, PYCD_ERR_ARG, "null lattice_q or site_energy");
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  const int64_t N = s->pb.n_sites;
  // the next trajectory starts from these sites: a running one is ended
  s->begun = false;
  PYCD_HIP(ctx, upload(s->qL.ptr, lattice_q, N, st));
  PYCD_HIP(ctx, upload(s->E.ptr, site_energy, N, st));
  PYCD_HIP(ctx, lattice_terms(s));
  PYCD_HIP(ctx, pack_wave_inputs(s));
  return PYCD_OK;
}

extern "C" void pycd_kmc_destroy(pycd_kmc* s) {
  if (!s) return;
  DeviceGuard guard(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  delete s;
}

extern "C" int pycd_kmc_begin(pycd_kmc* s, const int64_t* occupancy,
                              double ewald_neut) {
  if (!s || !occupancy) return PYCD_ERR_ARG;
  pycd_ctx* ctx = s->ctx;
  DeviceGuard guard(ctx->device);
  const pycd_kmc_problem& pb = s->pb;
  const int nc = pb.n_carriers;
  std::vector<int32_t> occ(nc);
  {
    // occupancy comes from the host API; validate every site.
    std::vector<int32_t> nb(size_t(pb.n_sites) * pb.max_z);
    PYCD_HIP(ctx, hipMemcpy(nb.data(), s->nbr.ptr, nb.size() * 4,
                            hipMemcpyDeviceToHost));
    for (int c = 0; c < nc; ++c) {
      const int64_t v = occupancy[c];
      if (v < 0 || v >= pb.n_sites)
        return fail(ctx, PYCD_ERR_ARG, "occupancy[%d]=%lld out of range", c,
                    (long long)v);
      for (int j = 0; j < pb.z; ++j)
        if (nb[size_t(v) * pb.max_z + j] < 0)
          return fail(ctx, PYCD_ERR_ARG,
                      "carrier %d sits on site %lld with fewer than z=%d hop "
                      "slots",
                      c, (long long)v, pb.z);
      occ[c] = int32_t(v);
    }
  }
  hipStream_t st = ctx->stream;
  const int64_t P = pb.n_path;
  PYCD_HIP(ctx, hipMemcpyAsync(s->occ.ptr, occ.data(), 4 * nc,
                               hipMemcpyHostToDevice, st));
  PYCD_HIP(ctx, hipMemsetAsync(s->disp.ptr, 0, 24 * nc, st));
  PYCD_HIP(ctx, hipMemsetAsync(s->base.ptr, 0, 24 * nc, st));
  PYCD_HIP(ctx, hipMemsetAsync(s->drift.ptr, 0, 24 * nc, st));
  PYCD_HIP(ctx, hipMemsetAsync(s->rows.ptr, 0, 24 * size_t(nc) * P, st));
  PYCD_HIP(ctx, hipMemsetAsync(s->epath.ptr, 0, 8 * size_t(P), st));
  PYCD_HIP(ctx, hipMemsetAsync(s->dpath.ptr, 0, 8 * size_t(P), st));
  carrier_energy_kernel<<<1, 256, 0, st>>>(s->A.ptr, s->phiL.ptr, s->cq.ptr,
                                          s->occ.ptr, nc, pb.n_sites,
                                          s->tmp.ptr);
  PYCD_HIP(ctx, hipGetLastError());
  double carrier_part = 0.0;
  PYCD_HIP(ctx, hipMemcpyAsync(&carrier_part, s->tmp.ptr, sizeof(double),
                               hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipStreamSynchronize(st));
  KmcScalars sc{};
  sc.t = 0.0;
  sc.energy = ewald_neut + (s->qLAqL + carrier_part);
  sc.start = 1;
  sc.hops = 0;
  sc.finished = P <= 1 ? 1 : 0;
  sc.no_select = 0;
  PYCD_HIP(ctx, hipMemcpyAsync(s->sc.ptr, &sc, sizeof(sc),
                               hipMemcpyHostToDevice, st));
  PYCD_HIP(ctx, hipMemcpyAsync(s->epath.ptr, &sc.energy, sizeof(double),
                               hipMemcpyHostToDevice, st));
  PYCD_HIP(ctx, hipStreamSynchronize(st));
  s->begun = true;
  return PYCD_OK;
}

namespace {

template <bool kCompact, int kNpl>
hipError_t launch_wave_t(size_t smem, hipStream_t st, const WaveArgs& wa,
                         const KmcDev* dv, int32_t* status) {
  if (smem > 65536) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&kmc_wave_kernel<kCompact, kNpl>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
    if (e != hipSuccess) return e;
  }
  kmc_wave_kernel<kCompact, kNpl><<<1, 64, smem, st>>>(wa, dv, status);
  return hipGetLastError();
}

template <bool kCompact>
hipError_t launch_wave_c(int npl, size_t smem, hipStream_t st,
                         const WaveArgs& wa, const KmcDev* dv,
                         int32_t* status) {
  switch (npl) {
    case 1:
      return launch_wave_t<kCompact, 1>(smem, st, wa, dv, status);
    case 2:
      return launch_wave_t<kCompact, 2>(smem, st, wa, dv, status);
    case 3:
      return launch_wave_t<kCompact, 3>(smem, st, wa, dv, status);
    case 4:
      return launch_wave_t<kCompact, 4>(smem, st, wa, dv, status);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_wave(bool compact, int npl, size_t smem, hipStream_t st,
                       const WaveArgs& wa, const KmcDev* dv,
                       int32_t* status) {
  return compact ? launch_wave_c<true>(npl, smem, st, wa, dv, status)
                 : launch_wave_c<false>(npl, smem, st, wa, dv, status);
}

}  // namespace

namespace pycd {

int hop_loop(pycd_kmc* s, const double* dU, double* LU, int32_t* status_dev,
             int64_t n_pairs, int64_t max_hops, double* dT, int32_t* dP,
             bool timed, int64_t* made);

// Device address of `scratch` on `device` (made current by the caller),
// resolved once per device and kept.
int kmc_static_scratch(int device, double** out) {
  static std::mutex mu;
  static std::vector<double*> addr;
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0) return PYCD_ERR_ARG;
  if (addr.size() <= size_t(device)) addr.resize(size_t(device) + 1, nullptr);
  if (!addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return PYCD_ERR_HIP;
    addr[device] = static_cast<double*>(p);
  }
  *out = addr[device];
  return PYCD_OK;
}

}  // namespace pycd

extern "C" int pycd_kmc_advance(pycd_kmc* s, const double* uniforms,
                                int64_t n_pairs, int64_t max_hops,
                                double* hop_time, int32_t* hop_proc,
                                int64_t* hops_out) {
  if (!s) return PYCD_ERR_ARG;
  pycd_ctx* ctx = s->ctx;
  if (!s->begun) return fail(ctx, PYCD_ERR_STATE, "pycd_kmc_begin not called");
  if (n_pairs < 0 || (n_pairs > 0 && !uniforms))
    return fail(ctx, PYCD_ERR_ARG, "bad uniform buffer");
  if (n_pairs > (int64_t(1) << 30))
    return fail(ctx, PYCD_ERR_ARG, "at most 2^30 uniform pairs per call");
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  const double* dU = uniforms;
  if (n_pairs > 0 && !is_device_ptr(uniforms)) {
    if (s->ubuf.count < size_t(2 * n_pairs))
      PYCD_HIP(ctx, s->ubuf.alloc(size_t(2 * n_pairs)));
    PYCD_HIP(ctx, hipMemcpyAsync(s->ubuf.ptr, uniforms, 16 * size_t(n_pairs),
                                 hipMemcpyHostToDevice, st));
    dU = s->ubuf.ptr;
  }
  double* dT = hop_time;
  int32_t* dP = hop_proc;
  const bool t_host = hop_time && !is_device_ptr(hop_time);
  const bool p_host = hop_proc && !is_device_ptr(hop_proc);
  if (t_host) {
    if (s->tbuf.count < size_t(n_pairs))
      PYCD_HIP(ctx, s->tbuf.alloc(size_t(std::max<int64_t>(n_pairs, 1))));
    dT = s->tbuf.ptr;
  }
  if (p_host) {
    if (s->pbuf.count < size_t(n_pairs))
      PYCD_HIP(ctx, s->pbuf.alloc(size_t(std::max<int64_t>(n_pairs, 1))));
    dP = s->pbuf.ptr;
  }
  if (s->lubuf.count < size_t(std::max<int64_t>(n_pairs, 1)))
    PYCD_HIP(ctx, s->lubuf.alloc(size_t(std::max<int64_t>(n_pairs, 1))));
  if (!s->ev0) {
    PYCD_HIP(ctx, hipEventCreate(&s->ev0));
    PYCD_HIP(ctx, hipEventCreate(&s->ev1));
  }
  int64_t made = 0;
  const int rc = hop_loop(s, dU, s->lubuf.ptr, s->status.ptr, n_pairs,
                          max_hops, dT, dP, true, &made);
  if (t_host && made > 0)
    PYCD_HIP(ctx, hipMemcpy(hop_time, dT, 8 * size_t(made),
                            hipMemcpyDeviceToHost));
  if (p_host && made > 0)
    PYCD_HIP(ctx, hipMemcpy(hop_proc, dP, 4 * size_t(made),
                            hipMemcpyDeviceToHost));
  if (hops_out) *hops_out = made;
  return rc;
}

namespace pycd {

// The hop loop of pycd_kmc_advance and pycd_kmc_advance_static: log(u2) of
// the chunk into LU, the kernel, its status into `status` (device) and back.
// Allocates nothing.  `timed`: HIP events around it (created beforehand).
int hop_loop(pycd_kmc* s, const double* dU, double* LU, int32_t* status_dev,
             int64_t n_pairs, int64_t max_hops, double* dT, int32_t* dP,
             bool timed, int64_t* made) {
  pycd_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const pycd_kmc_problem& pb = s->pb;
  KmcDev d;
  d.A = s->A.ptr;
  d.Ac = s->Ac.ptr;
  d.pk = s->pk.ptr;
  d.nbr_pk = s->nbr_pk.ptr;
  d.gp = geo_pk(s->compact ? s->geo : Geo{1, 1, 1, 1});
  d.phiL = s->phiL.ptr;
  d.dphi = s->dphi.ptr;
  d.nbr = s->nbr.ptr;
  d.vec = s->vec.ptr;
  d.lam = s->lam.ptr;
  d.vab = s->vab.ptr;
  d.tself = s->tself.ptr;
  d.E = s->E.ptr;
  d.cq = s->cq.ptr;
  d.N = pb.n_sites;
  d.nc = pb.n_carriers;
  d.z = pb.z;
  d.max_z = pb.max_z;
  d.field_on = (pb.field[0] != 0.0 || pb.field[1] != 0.0 || pb.field[2] != 0.0);
  for (int a = 0; a < 3; ++a) d.field[a] = pb.field[a];
  d.nu = pb.nu;
  d.kT = pb.kT;
  d.nkt = -1.0 / pb.kT;
  d.dt = pb.time_interval;
  d.P = pb.n_path;
  d.occ = s->occ.ptr;
  d.disp = s->disp.ptr;
  d.base = s->base.ptr;
  d.drift = s->drift.ptr;
  d.rows = s->rows.ptr;
  d.epath = s->epath.ptr;
  d.dpath = s->dpath.ptr;
  d.sc = s->sc.ptr;
  if (timed) PYCD_HIP(ctx, hipEventRecord(s->ev0, st));
  reset_status<<<1, 1, 0, st>>>(status_dev);
  if (n_pairs > 0)
    log_u2_kernel<<<int(std::min<int64_t>((n_pairs + 255) / 256, 4096)), 256,
                    0, st>>>(dU, n_pairs, LU);
  const bool wave = s->variant == PYCD_KMC_WAVE ||
                    (s->variant == PYCD_KMC_AUTO && s->wave_ok);
  if (wave) {
    WaveArgs wa;
    wa.tab = reinterpret_cast<const char*>(s->compact ? s->Ac.ptr : s->A.ptr);
    wa.row_bytes = uint32_t(pb.n_sites) * uint32_t(sizeof(double));
    wa.gp = d.gp;
    wa.rec = s->rec.ptr;
    wa.pe = s->pe.ptr;
    wa.nbr2 = s->nbr2.ptr;
    wa.rows = d.rows;
    wa.epath = d.epath;
    wa.dpath = d.dpath;
    wa.U = dU;
    wa.LU = LU;
    wa.hop_time = dT;
    wa.hop_proc = dP;
    wa.nc = pb.n_carriers;
    wa.z = pb.z;
    wa.mz = pb.max_z;
    wa.field_on = d.field_on;
    wa.uq = s->uniform_q ? 1 : 0;
    wa.q0 = s->q0;
    wa.nu = pb.nu;
    wa.nkt = d.nkt;
    wa.dt = pb.time_interval;
    wa.P = pb.n_path;
    wa.n_pairs = n_pairs;
    wa.max_hops = max_hops;
    PYCD_HIP(ctx, hipMemcpyAsync(s->dev.ptr, &d, sizeof(KmcDev),
                                 hipMemcpyHostToDevice, st));
    PYCD_HIP(ctx, launch_wave(s->compact, s->wave_npl, s->wave_smem, st, wa,
                              s->dev.ptr, status_dev));
  } else if (s->compact) {
    kmc_plain_kernel<true><<<1, kThreads, s->smem, st>>>(
        d, dU, LU, n_pairs, max_hops, dT, dP, status_dev);
  } else {
    kmc_plain_kernel<false><<<1, kThreads, s->smem, st>>>(
        d, dU, LU, n_pairs, max_hops, dT, dP, status_dev);
  }
  PYCD_HIP(ctx, hipGetLastError());
  if (timed) PYCD_HIP(ctx, hipEventRecord(s->ev1, st));
  int32_t status[2] = {-1, 0};
  PYCD_HIP(ctx, hipMemcpyAsync(status, status_dev, sizeof(status),
                               hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipStreamSynchronize(st));
  if (timed) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    s->last_ms = ms;
  }
  *made = status[1];
  if (status[0] < 0) return fail(ctx, PYCD_ERR_STATE, "kernel did not report");
  return status[0] == 1 ? PYCD_NEED_UNIFORMS : PYCD_OK;
}

}  // namespace pycd

extern "C" int pycd_kmc_advance_static(pycd_kmc* s, const double* uniforms,
                                       int64_t n_pairs, int64_t max_hops,
                                       double* hop_time, int32_t* hop_proc,
                                       int64_t* hops_out) {
  if (!s) return PYCD_ERR_ARG;
  pycd_ctx* ctx = s->ctx;
  if (!s->begun) return fail(ctx, PYCD_ERR_STATE, "pycd_kmc_begin not called");
  if (n_pairs < 0 || (n_pairs > 0 && !uniforms))
    return fail(ctx, PYCD_ERR_ARG, "bad uniform buffer");
  if (n_pairs > kStaticPairs)
    return fail(ctx, PYCD_ERR_ARG,
                "pycd_kmc_advance_static takes at most %lld uniform pairs "
                "per call (its log(u2) scratch is a fixed device array); got "
                "%lld",
                (long long)kStaticPairs, (long long)n_pairs);
  if (!ctx->static_scratch)
    return fail(ctx, PYCD_ERR_STATE, "static scratch of the device unresolved");
  DeviceGuard guard(ctx->device);
  if ((n_pairs > 0 && !is_device_ptr(uniforms)) ||
      (hop_time && !is_device_ptr(hop_time)) ||
      (hop_proc && !is_device_ptr(hop_proc)))
    return fail(ctx, PYCD_ERR_ARG,
                "pycd_kmc_advance_static takes device buffers only");
  double* LU = ctx->static_scratch;
  int32_t* status = reinterpret_cast<int32_t*>(LU + kStaticPairs);
  int64_t made = 0;
  const int rc = hop_loop(s, uniforms, LU, status, n_pairs, max_hops,
                          hop_time, hop_proc, false, &made);
  if (hops_out) *hops_out = made;
  return rc;
}

extern "C" int pycd_kmc_read(pycd_kmc* s, double* unwrapped, double* energy,
                             double* delg_0, double* drift, int64_t* occupancy,
                             double scalars[4]) {
  if (!s) return PYCD_ERR_ARG;
  pycd_ctx* ctx = s->ctx;
  DeviceGuard guard(ctx->device);
  hipStream_t st = ctx->stream;
  const pycd_kmc_problem& pb = s->pb;
  const int nc = pb.n_carriers;
  const int64_t P = pb.n_path;
  if (unwrapped)
    PYCD_HIP(ctx, download(unwrapped, s->rows.ptr, size_t(P) * 3 * nc, st));
  if (energy) PYCD_HIP(ctx, download(energy, s->epath.ptr, size_t(P), st));
  if (delg_0) PYCD_HIP(ctx, download(delg_0, s->dpath.ptr, size_t(P), st));
  if (drift) PYCD_HIP(ctx, download(drift, s->drift.ptr, 3 * size_t(nc), st));
  KmcScalars sc{};
  PYCD_HIP(ctx, hipMemcpyAsync(&sc, s->sc.ptr, sizeof(sc),
                               hipMemcpyDeviceToHost, st));
  std::vector<int32_t> occ(nc);
  PYCD_HIP(ctx, hipMemcpyAsync(occ.data(), s->occ.ptr, 4 * nc,
                               hipMemcpyDeviceToHost, st));
  PYCD_HIP(ctx, hipStreamSynchronize(st));
  if (occupancy)
    for (int c = 0; c < nc; ++c) occupancy[c] = occ[c];
  if (scalars) {
    scalars[0] = sc.t;
    scalars[1] = sc.energy;
    scalars[2] = double(sc.start);
    scalars[3] = double(sc.hops);
  }
  return PYCD_OK;
}

extern "C" double pycd_kmc_last_ms(const pycd_kmc* s) {
  return s ? s->last_ms : -1.0;
}

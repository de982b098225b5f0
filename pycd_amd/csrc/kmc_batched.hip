// Batched KMC: thousands of independent Run.do_kmc_steps trajectories
// (PyCD/core.py:2668-2861) of uneven supercell size and carrier count in
// one persistent launch on gfx950.
//
// Scheduling: one wavefront (64 lanes, one workgroup) owns one trajectory at
// a time and takes the next from a global atomic queue when it finishes, in
// the order given by the host (longest expected work first).  The queue has
// an exit condition every wave reaches (index >= n_traj), so the grid drains.
//
// Per-hop algebra is the plain path's (q.A[x,:] = phi_L[x] + psi[x]), but
// psi is maintained incrementally: when carrier m hops s -> n, every other
// carrier's slot values change by q_m (A[n, x] - A[s, x]) (two gathers per
// slot) and carrier m's own z+1 slots are recomputed from scratch
// (n_c gathers each).  That is 3 (1 + z) n_c table reads per hop.
//
// Random numbers: each trajectory carries its CPython MT19937 state (624
// words + position) in LDS and draws 4 words per hop exactly like
// random.random() (u1 selects, u2 advances the clock), so a batched
// trajectory consumes the same stream as the plain path would.
//
// Rates: FP64 (bit-faithful mode) or FP32 (throughput mode); the prefix
// sum of k / k_tot, the selection and the clock are FP64 in both.
#include "common.hpp"
#include "mt19937.hpp"

#include <algorithm>
#include <cmath>

namespace pycd {

struct SlotRec {      // one hop slot of one site; 64 bytes
  int32_t nbr;        // target site, -1 if the slot does not exist
  int32_t nbr_row;    // hopping-site row of the target (difference table)
  double lam;         // lambda (Hartree)
  double vab;         // V_ab (Hartree)
  double tself;       // A[s,s] - A[s,n]
  double vec[3];      // hop vector (bohr)
  double pad2;
};
static_assert(sizeof(SlotRec) == 64, "slot record is one 64-byte line");

struct FamDev {
  const double* A64;
  const float* A32;
  const double* D64;      // (hop row, slot) x N: A[nbr(s,j), :] - A[s, :]
  const float* D32;
  const int32_t* hop_row; // site -> compact hopping-site row, -1 otherwise
  const double* phiL;
  const double* E;
  const SlotRec* slots;
  int64_t N;
  int32_t z, mz;
  double qLAqL;
};

struct TrajDev {
  int32_t fam, nc;
  double q;           // carrier charge
  double field[3];    // au
  double kT;          // Hartree
  double dt;          // path-step width (au time)
  double ewald_neut;
  int64_t P;          // path rows
  int64_t max_hops;   // <= 0: run until the path is full
  int64_t row_off;    // offset (doubles) of this trajectory's path rows
  int64_t occ_off;    // offset of its initial occupancy
};

struct TrajOut {
  double t;
  double energy;
  int64_t hops;
  int64_t start;
  int32_t no_select;
  int32_t done;      // 1: the trajectory reached max_hops or filled its path
  int32_t started;   // 1: state arrays hold a run in progress
  int32_t pad0;
  double pad1;
};

struct BatchArgs {
  const FamDev* fams;
  const TrajDev* trajs;
  const int32_t* order;
  const int64_t* occ0;
  const uint32_t* mt0;      // n_traj x 625
  double* rows;             // all path rows
  double* drift;            // n_traj x 64 x 3 (nc <= kMaxCarriers)
  TrajOut* out;
  uint32_t* counter;
  // resumable state (slices): MT words, occupancy, displacement/base/drift
  uint32_t* mts;            // n_traj x 625
  int32_t* st_occ;          // n_traj x kMaxCarriers
  double* st_vec;           // n_traj x 3 x (3 kMaxCarriers)
  double* st_psi;           // n_traj x kPsiStride
  int64_t slice_hops;       // <= 0: run each trajectory to its end
  int64_t refresh;          // recompute every psi from the table each
                            // `refresh` hops of a trajectory (<= 0: never)
  int32_t n_traj;
  int32_t nc_max;
  int32_t z_max;
  double nu;
};

constexpr int kMaxCarriers = 64;
constexpr int kMaxPer = 8;    // processes per lane: n_c z <= 512
constexpr int kPsiStride = kMaxCarriers * 8;  // n_c (1 + z), z <= 7
constexpr int kVecStride = 3 * 3 * kMaxCarriers;
constexpr double kEulerE = 2.718281828459045;

// Values every lane holds identically, moved to scalar registers.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uni(double v) {
  const uint64_t b = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(b));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(b >> 32));
  return __longlong_as_double((int64_t(hi) << 32) | lo);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double wave_excl_scan(double v, int lane) {
  double incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  const double excl = __shfl_up(incl, 1, 64);
  return lane == 0 ? 0.0 : excl;
}

template <typename TT>
__device__ inline double tab(const FamDev& f, int64_t idx);
template <>
__device__ inline double tab<double>(const FamDev& f, int64_t idx) {
  return f.A64[idx];
}
template <>
__device__ inline double tab<float>(const FamDev& f, int64_t idx) {
  return double(f.A32[idx]);
}

// Row pointers (uniform across the wavefront) for the gathers.
template <typename TT>
__device__ inline const TT* trow(const FamDev& f, int64_t row);
template <>
__device__ inline const double* trow<double>(const FamDev& f, int64_t row) {
  return f.A64 + row * f.N;
}
template <>
__device__ inline const float* trow<float>(const FamDev& f, int64_t row) {
  return f.A32 + row * f.N;
}
template <typename TT>
__device__ inline const TT* drow_ptr(const FamDev& f, int64_t row);
template <>
__device__ inline const double* drow_ptr<double>(const FamDev& f, int64_t row) {
  return f.D64 + row * f.N;
}
template <>
__device__ inline const float* drow_ptr<float>(const FamDev& f, int64_t row) {
  return f.D32 + row * f.N;
}

template <typename TT>
__device__ inline double dtab(const FamDev& f, int64_t idx);
template <>
__device__ inline double dtab<double>(const FamDev& f, int64_t idx) {
  return f.D64[idx];
}
template <>
__device__ inline double dtab<float>(const FamDev& f, int64_t idx) {
  return double(f.D32[idx]);
}

// psi[r] = sum_c q A[sites[r], occ[c]] for every slot r (lanes over slots).
template <typename TT>
__device__ inline void psi_from_table(const FamDev& f, double q, int R, int nc,
                                      const int32_t* sites,
                                      const int32_t* occ, double* psi,
                                      int lane) {
  for (int r = lane; r < R; r += 64) {
    const int64_t row = int64_t(sites[r]) * f.N;
    double acc = 0.0;
#pragma unroll 4
    for (int c2 = 0; c2 < nc; ++c2) acc = fma(q, tab<TT>(f, row + occ[c2]), acc);
    psi[r] = acc;
  }
}

// Shared-memory layout of one wavefront (sizes from the batch's maxima:
// pmax = nc_max z_max processes, rmax = nc_max (1 + z_max) slots):
//   mt     624 words (2560 B)        MT19937 block
//   psi    rmax doubles              potential of the carriers at each slot
//   pcd    pmax x kPcD doubles       static part of each process's dG0
//   pcf    pmax x 2 floats           (FP32 rates) lambda, V_ab + field shift
//   sites  rmax ints                 slot sites (carrier site, its targets)
//   nn     pmax x z_max ints         hop targets of each process's target
//   occ, ohr  nc_max ints each       carrier sites and their hop rows
template <bool kF32Rate>
struct PcLayout {
  static constexpr int kPcD = kF32Rate ? 1 : 6;
  static constexpr int kPcF = kF32Rate ? 2 : 0;
};

__host__ __device__ inline size_t batch_smem_bytes(int nc_max, int z_max,
                                                  bool f32rate) {
  const size_t rmax = size_t(nc_max) * (1 + z_max);
  const size_t pmax = size_t(nc_max) * z_max;
  const size_t pcd = f32rate ? 1 : 6, pcf = f32rate ? 2 : 0;
  return 2560 + 8 * rmax + 8 * pmax * pcd + 4 * pmax * pcf + 4 * rmax +
         4 * pmax * z_max + 8 * size_t(nc_max);
}

// kPer: processes per lane (n_c z <= 64 kPer); kZp: slots per carrier
// (z + 1 <= kZp).  Instantiated for the batch's maxima so that the per-lane
// arrays stay small and several wavefronts fit per SIMD.
#ifndef PYCD_BATCH_WAVES_PER_EU
#define PYCD_BATCH_WAVES_PER_EU 2
#endif
template <typename TT, bool kF32Rate, int kPer, int kZp>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(PYCD_BATCH_WAVES_PER_EU)))
kmc_batched_kernel(BatchArgs a) {
  using L = PcLayout<kF32Rate>;
  extern __shared__ unsigned char smem_raw[];
  const int lane = threadIdx.x;
  const int zmax = a.z_max;
  const int rmax = a.nc_max * (1 + zmax);
  const int pmax = a.nc_max * zmax;
  uint32_t* mt = reinterpret_cast<uint32_t*>(smem_raw);
  double* psi = reinterpret_cast<double*>(smem_raw + 2560);
  double* pcd = psi + rmax;
  float* pcf = reinterpret_cast<float*>(pcd + size_t(pmax) * L::kPcD);
  int32_t* sites = reinterpret_cast<int32_t*>(pcf + size_t(pmax) * L::kPcF);
  int32_t* nn = sites + rmax;
  int32_t* occ = nn + size_t(pmax) * zmax;
  int32_t* ohr = occ + a.nc_max;

  for (;;) {
    uint32_t ti = 0;
    if (lane == 0) ti = atomicAdd(a.counter, 1u);
    ti = uint32_t(uni(int(__shfl(ti, 0, 64))));
    if (ti >= uint32_t(a.n_traj)) break;
    const int tr = uni(a.order[ti]);
    const TrajOut o0 = a.out[tr];
    if (uni(o0.done)) continue;
    const bool resume = uni(o0.started) != 0;
    const TrajDev td = a.trajs[tr];
    const FamDev f = a.fams[uni(td.fam)];
    const int nc = td.nc, z = f.z, zp = z + 1;
    const int R = nc * zp, nproc = nc * z;
    const int per = (nproc + 63) / 64;
    const int64_t N = f.N;
    const double q = td.q;
    double* rows = a.rows + td.row_off;
    const int w = 3 * nc;
    const bool field_on =
        td.field[0] != 0.0 || td.field[1] != 0.0 || td.field[2] != 0.0;

    // per-trajectory state in global memory: each entry k of disp, base and
    // drift is read and written only by lane k % 64
    uint32_t* mts = a.mts + int64_t(tr) * 625;
    int32_t* socc = a.st_occ + int64_t(tr) * kMaxCarriers;
    double* svec = a.st_vec + int64_t(tr) * kVecStride;
    double* sdisp = svec;
    double* sbase = svec + 3 * kMaxCarriers;
    double* sdrift = svec + 6 * kMaxCarriers;
    double* spsi = a.st_psi + int64_t(tr) * kPsiStride;

    // ---- load or start the state
    const uint32_t* m0 = resume ? mts : a.mt0 + int64_t(tr) * 625;
    for (int k = lane; k < 624; k += 64) mt[k] = m0[k];
    int mti = int(m0[624]);
    for (int c = lane; c < nc; c += 64) {
      occ[c] = resume ? socc[c] : int32_t(a.occ0[td.occ_off + c]);
      ohr[c] = f.hop_row[occ[c]];
    }
    if (!resume)
      for (int k = lane; k < w; k += 64) {
        sdisp[k] = 0.0;
        sbase[k] = 0.0;
        sdrift[k] = 0.0;
        rows[k] = 0.0;  // path row 0: the initial positions
      }
    __syncthreads();
    for (int r = lane; r < R; r += 64) {
      const int c = r / zp, k = r - c * zp;
      const int s = occ[c];
      sites[r] = k == 0 ? s : f.slots[int64_t(s) * f.mz + k - 1].nbr;
    }
    __syncthreads();
    // process cache and second neighbours of every process
    for (int p = lane; p < nproc; p += 64) {
      const int c = p / z, j = p - c * z;
      const int s = sites[c * zp];
      const SlotRec& rec = f.slots[int64_t(s) * f.mz + j];
      const int n = rec.nbr;
      const double dphi = f.phiL[n] - f.phiL[s];
      const double tq = q * rec.tself;
      const double dE = f.E[n] - f.E[s];
      const double shift =
          field_on ? 0.5 * (td.field[0] * rec.vec[0] + td.field[1] * rec.vec[1] +
                            td.field[2] * rec.vec[2])
                   : 0.0;
      if (kF32Rate) {
        pcd[p] = 2 * q * (dphi + tq) + dE;
        pcf[2 * p] = float(rec.lam);
        pcf[2 * p + 1] = float(rec.vab) + float(shift);
      } else {
        double* e = pcd + 6 * p;
        e[0] = dphi;
        e[1] = tq;
        e[2] = dE;
        e[3] = rec.lam;
        e[4] = rec.vab;
        e[5] = shift;
      }
    }
    for (int pk = lane; pk < nproc * z; pk += 64) {
      const int p = pk / z, k = pk - p * z;
      const int c = p / z, j = p - c * z;
      const int t = sites[c * zp + 1 + j];
      nn[p * zmax + k] = f.slots[int64_t(t) * f.mz + k].nbr;
    }
    double t_clk, energy;
    int64_t start, hops;
    int no_select;
    if (resume) {
      for (int r = lane; r < R; r += 64) psi[r] = spsi[r];
      t_clk = o0.t;
      energy = o0.energy;
      start = o0.start;
      hops = o0.hops;
      no_select = o0.no_select;
    } else {
      psi_from_table<TT>(f, q, R, nc, sites, occ, psi, lane);
      // initial energy: ewald_neut + q_L A q_L + sum_c q (2 phi_L + psi_self)
      double e_loc = 0.0;
      for (int c = lane; c < nc; c += 64) {
        const int64_t row = int64_t(occ[c]) * N;
        double inner = 0.0;
        for (int c2 = 0; c2 < nc; ++c2)
          inner = fma(q, tab<TT>(f, row + occ[c2]), inner);
        e_loc += q * (2.0 * f.phiL[occ[c]] + inner);
      }
      energy = td.ewald_neut + (f.qLAqL + wave_sum(e_loc));
      t_clk = 0.0;
      start = 1;
      hops = 0;
      no_select = 0;
    }
    __syncthreads();
    int64_t slice_done = 0;
    bool finished = false;
    const float kTf = float(td.kT);
    const float nuf = float(a.nu);

    for (;;) {
      if (td.max_hops > 0 && hops >= td.max_hops) {
        finished = true;
        break;
      }
      if (a.slice_hops > 0 && slice_done >= a.slice_hops) break;
      ++slice_done;
      if (a.refresh > 0 && hops > 0 && hops % a.refresh == 0) {
        // bound the rounding carried by the incremental updates
        psi_from_table<TT>(f, q, R, nc, sites, occ, psi, lane);
        __syncthreads();
      }
      // ---- rates from shared memory only
      double kv[kPer], gv[kPer];
      double ksum = 0.0;
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        kv[i] = 0.0;
        gv[i] = 0.0;
        const int p = lane * per + i;
        if (i < per && p < nproc) {
          const int c = p / z;
          const int j = p - c * z;
          const double dpsi = psi[c * zp + 1 + j] - psi[c * zp];
          double g0, k;
          if (kF32Rate) {
            g0 = pcd[p] + 2 * q * dpsi;
            const float lam = pcf[2 * p];
            const float gf = float(g0);
            const float gs = (lam + gf) * (lam + gf) / (4.f * lam) - pcf[2 * p + 1];
            k = double(nuf * __expf(-gs / kTf));
          } else {
            const double* e = pcd + 6 * p;
            g0 = 2 * q * ((e[0] + dpsi) + e[1]) + e[2];
            const double lam = e[3];
            const double gs =
                (((lam + g0) * (lam + g0) / (4 * lam)) - e[4]) - e[5];
            k = a.nu * pow(kEulerE, -gs / td.kT);
          }
          kv[i] = k;
          gv[i] = g0;
          ksum += k;
        }
      }
      const double k_tot = uni(wave_sum(ksum));
      // ---- two uniforms, as random.random() draws them
      uint32_t wd[4];
      mt_next4(mt, mti, lane, wd);
      const double u1 = mt_uniform(wd[0], wd[1]);
      const double u2 = mt_uniform(wd[2], wd[3]);
      // ---- FP64 prefix of k / k_tot and selection
      double xs = 0.0;
#pragma unroll
      for (int i = 0; i < kPer; ++i) xs += kv[i] / k_tot;
      double cum = wave_excl_scan(xs, lane);
      int found = -1;
      double g_found = 0.0, k_found = 0.0;
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        const int p = lane * per + i;
        if (i < per && p < nproc) {
          cum += kv[i] / k_tot;
          if (found < 0 && cum > u1) {
            found = p;
            g_found = gv[i];
            k_found = kv[i];
          }
        }
      }
      const unsigned long long mask = __ballot(found >= 0);
      int src;
      if (mask) {
        src = __ffsll(static_cast<long long>(mask)) - 1;
      } else {
        // rounding left every cum <= u1: take the last process
        src = (nproc - 1) / per;
        if (lane == src) {
          found = nproc - 1;
          const int i = nproc - 1 - lane * per;
#pragma unroll
          for (int ii = 0; ii < kPer; ++ii)
            if (ii == i) {
              g_found = gv[ii];
              k_found = kv[ii];
            }
        }
        ++no_select;
      }
      const int sel = uni(__shfl(found, src, 64));
      const double g = uni(__shfl(g_found, src, 64));
      const double ksel = uni(__shfl(k_found, src, 64));
      // ---- clock
      int64_t end;
      if (u2 > 0.0) {
        t_clk -= log(u2) / k_tot;
        const double qq = t_clk / td.dt;
        end = qq < 9.0e18 ? int64_t(qq) : td.P;
      } else {
        t_clk = INFINITY;
        end = td.P;
      }
      energy += g;
      // ---- one round of independent loads for the whole update
      const int m = sel / z;
      const int j = sel - m * z;
      const int s_old = uni(sites[m * zp]);
      const int s_new = uni(sites[m * zp + 1 + j]);
      const TT* drp = drow_ptr<TT>(f, int64_t(uni(ohr[m])) * z + j);
      int nsel[kZp - 1];  // targets of the new site, from shared memory
#pragma unroll
      for (int k = 0; k < kZp - 1; ++k)
        nsel[k] = k < z ? uni(nn[sel * zmax + k]) : -1;
      constexpr int kRPer = kMaxCarriers * kZp / 64;
      double dv[kRPer];
#pragma unroll
      for (int i = 0; i < kRPer; ++i) {
        const int r = lane + 64 * i;
        dv[i] = r < R ? double(drp[uint32_t(sites[r])]) : 0.0;
      }
      // fresh sums of the new slots (not the one that returns to s_old):
      // one carrier per lane (n_c <= 64)
      double part[kZp - 1];
      {
        const int o = lane < nc ? (lane == m ? s_new : occ[lane]) : -1;
#pragma unroll
        for (int k = 0; k < kZp - 1; ++k) {
          part[k] = 0.0;
          if (k < z && o >= 0 && nsel[k] != s_old)
            part[k] = q * double(trow<TT>(f, nsel[k])[uint32_t(o)]);
        }
      }
      // slot records of the new site (lanes < z): next process cache
      SlotRec rec_new;
      double phin = 0.0, en = 0.0, phis = 0.0, es = 0.0;
      if (lane < z) {
        rec_new = f.slots[int64_t(s_new) * f.mz + lane];
        phin = f.phiL[rec_new.nbr];
        en = f.E[rec_new.nbr];
        phis = f.phiL[s_new];
        es = f.E[s_new];
      }
      // second neighbours of the new processes (lanes < z z)
      int nn_new = -1;
      if (lane < z * z) {
        const int jj = lane / z, kk = lane - jj * z;
        nn_new = f.slots[int64_t(nsel[jj < kZp - 1 ? jj : 0]) * f.mz + kk].nbr;
      }
      const int nrow = uni(f.slots[int64_t(s_old) * f.mz + j].nbr_row);
      // displacement by the owner lanes of carrier m's three components
      double vec_own = 0.0;
      int own_k = -1;  // the component of carrier m this lane owns, if any
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (lane == (3 * m + cc) % 64) own_k = 3 * m + cc;
      const bool owner = own_k >= 0;
      if (owner)
        vec_own = f.slots[int64_t(s_old) * f.mz + j].vec[own_k - 3 * m];
      // ---- apply
#pragma unroll
      for (int i = 0; i < kRPer; ++i) {
        const int r = lane + 64 * i;
        if (r < R) psi[r] += q * dv[i];
      }
      if (owner) {
        sdisp[own_k] += vec_own;
        if (field_on) sdrift[own_k] += vec_own * ksel;
      }
      __syncthreads();
      const double v_new = uni(psi[m * zp + 1 + j]);
      const double v_old = uni(psi[m * zp]);
      double fresh_v[kZp - 1];
#pragma unroll
      for (int k = 0; k < kZp - 1; ++k)
        fresh_v[k] = (k < z && nsel[k] != s_old) ? wave_sum(part[k]) : v_old;
      __syncthreads();
      if (lane == 0) {
        occ[m] = s_new;
        ohr[m] = nrow;
        sites[m * zp] = s_new;
        psi[m * zp] = v_new;
#pragma unroll
        for (int k = 0; k < kZp - 1; ++k)
          if (k < z) {
            sites[m * zp + 1 + k] = nsel[k];
            psi[m * zp + 1 + k] = fresh_v[k];
          }
      }
      if (lane < z) {
        const int pn = m * z + lane;
        const double dphi = phin - phis;
        const double tq = q * rec_new.tself;
        const double dE = en - es;
        const double shift =
            field_on ? 0.5 * (td.field[0] * rec_new.vec[0] +
                              td.field[1] * rec_new.vec[1] +
                              td.field[2] * rec_new.vec[2])
                     : 0.0;
        if (kF32Rate) {
          pcd[pn] = 2 * q * (dphi + tq) + dE;
          pcf[2 * pn] = float(rec_new.lam);
          pcf[2 * pn + 1] = float(rec_new.vab) + float(shift);
        } else {
          double* e = pcd + 6 * pn;
          e[0] = dphi;
          e[1] = tq;
          e[2] = dE;
          e[3] = rec_new.lam;
          e[4] = rec_new.vab;
          e[5] = shift;
        }
      }
      if (lane < z * z) {
        const int jj = lane / z, kk = lane - jj * z;
        nn[(m * z + jj) * zmax + kk] = nn_new;
      }
      // ---- path rows [start, end), written by the owner lane of each entry
      ++hops;
      bool full = false;
      if (end >= start + 1) {
        if (end >= td.P) end = td.P;
        for (int k = lane; k < w; k += 64) {
          const double v = sbase[k] + sdisp[k];
          for (int64_t r = start; r < end; ++r) rows[r * w + k] = v;
          sbase[k] = v;
          sdisp[k] = 0.0;
        }
        start = end;
        full = end >= td.P;
      }
      __syncthreads();
      if (full) {
        finished = true;
        break;
      }
    }
    // ---- outputs and resumable state
    for (int k = lane; k < w; k += 64)
      a.drift[int64_t(tr) * 3 * kMaxCarriers + k] = sdrift[k];
    for (int k = lane; k < 624; k += 64) mts[k] = mt[k];
    for (int c = lane; c < nc; c += 64) socc[c] = occ[c];
    for (int r = lane; r < R; r += 64) spsi[r] = psi[r];
    if (lane == 0) {
      mts[624] = uint32_t(mti);
      TrajOut o;
      o.t = t_clk;
      o.energy = energy;
      o.hops = hops;
      o.start = start;
      o.no_select = no_select;
      o.done = finished ? 1 : 0;
      o.started = 1;
      o.pad0 = 0;
      o.pad1 = 0.0;
      a.out[tr] = o;
    }
    __syncthreads();
  }
}

// Squared displacement per trajectory, averaged over its carriers:
//   msd[t, tau] = mean_c mean_{t0 < P_t - tau} |r_c(t0 + tau) - r_c(t0)|^2
// One workgroup per (trajectory, block of 8 lags).
struct MsdTraj {
  int64_t row_off;
  int64_t P;
  int32_t nc;
  int32_t traj;     // index into the trajectory results (rows written)
  int64_t out_off;
  int64_t M;
};

__global__ void __launch_bounds__(256)
msd_batched_kernel(const double* __restrict__ rows,
                   const MsdTraj* __restrict__ tr,
                   const TrajOut* __restrict__ outs, int n_traj, int nlb_max,
                   double scale, double* __restrict__ out) {
  const int t = blockIdx.x / nlb_max;
  const int lb = blockIdx.x - t * nlb_max;
  if (t >= n_traj) return;
  MsdTraj d = tr[t];
  // only the rows written by the run: [0, start) (all P when the path filled)
  const int64_t written = outs[d.traj].start;
  const int64_t M_all = d.M;
  d.P = written < 1 ? 1 : (written < d.P ? written : d.P);
  const int64_t tau0 = int64_t(lb) * 8;
  if (tau0 >= M_all) return;
  const int w = 3 * d.nc;
  const double* base = rows + d.row_off;
  double acc[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) acc[l] = 0.0;
  // per lag the carriers are summed with the 1/(P - tau) weight applied at
  // the end (same for every carrier of the trajectory)
  for (int64_t idx = threadIdx.x; idx < d.P * d.nc; idx += blockDim.x) {
    const int64_t t0 = idx / d.nc;
    const int c = int(idx - t0 * d.nc);
    const double* p0 = base + t0 * w + 3 * c;
    const double x0 = p0[0] * scale, y0 = p0[1] * scale, z0 = p0[2] * scale;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const int64_t tau = tau0 + l;
      if (tau >= 1 && tau < d.M && t0 + tau < d.P) {
        const double* p1 = base + (t0 + tau) * w + 3 * c;
        const double dx = p1[0] * scale - x0, dy = p1[1] * scale - y0,
                     dz = p1[2] * scale - z0;
        acc[l] += dx * dx + dy * dy + dz * dz;
      }
    }
  }
  __shared__ double part[8][256];
#pragma unroll
  for (int l = 0; l < 8; ++l) part[l][threadIdx.x] = acc[l];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int l = 0; l < 8; ++l) part[l][threadIdx.x] += part[l][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 8) {
    const int64_t tau = tau0 + threadIdx.x;
    if (tau < M_all) {
      double v = 0.0;
      if (tau >= 1 && tau < d.M && tau < d.P)
        v = part[threadIdx.x][0] / double(d.P - tau) / double(d.nc);
      out[d.out_off + tau] = v;
    }
  }
}

__global__ void fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void reset_outputs(TrajOut* o, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    o[i].done = 0;
    o[i].started = 0;
  }
}

__global__ void fill_f64(double* p, int64_t n, double v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Family preparation kernels.
__global__ void __launch_bounds__(256)
row_dot_f64(const double* __restrict__ A, const double* __restrict__ q,
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

__global__ void diff_table(const double* __restrict__ A,
                           const SlotRec* __restrict__ slots,
                           const int32_t* __restrict__ hop_sites, int64_t n_hop,
                           int z, int mz, int64_t N, double* __restrict__ D) {
  const int64_t total = n_hop * z * N;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t x = i % N;
    const int64_t hj = i / N;
    const int64_t h = hj / z;
    const int j = int(hj - h * z);
    const int64_t s = hop_sites[h];
    const int n = slots[s * mz + j].nbr;
    D[i] = A[int64_t(n) * N + x] - A[s * N + x];
  }
}

__global__ void to_f32(const double* __restrict__ a, float* __restrict__ b,
                       int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = float(a[i]);
}

__global__ void slot_self_terms(const double* __restrict__ A, int64_t N,
                                SlotRec* __restrict__ slots, int mz) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= N * mz) return;
  const int64_t s = idx / mz;
  const int n = slots[idx].nbr;
  slots[idx].tself = n >= 0 ? A[s * N + s] - A[s * N + n] : 0.0;
}

__global__ void dot_f64(const double* __restrict__ a,
                        const double* __restrict__ b, int64_t n,
                        double* __restrict__ out) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) acc = fma(a[k], b[k], acc);
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

}  // namespace pycd

// ------------------------------------------------------------------- host --
using namespace pycd;

struct BatchFamily {
  DevBuf<double> A64, phiL, E, qL, tmp, D64;
  DevBuf<float> A32, D32;
  DevBuf<int32_t> hop_row, hop_sites;
  int64_t n_hop = 0;
  DevBuf<SlotRec> slots;
  int64_t N = 0;
  int z = 0, mz = 0;
  double qLAqL = 0.0;
  std::vector<int32_t> nbr_host;  // for host-side validation
};

struct pycd_batch {
  pycd_ctx* ctx = nullptr;
  std::vector<BatchFamily*> fams;
  DevBuf<FamDev> d_fams;
  int table_prec = 0;  // 0 FP64, 1 FP32
  int64_t refresh = 1024;
  int rate_prec = 0;   // 0 FP64, 1 FP32
  double nu = 0.0;
  // per-run buffers
  DevBuf<TrajDev> d_trajs;
  DevBuf<int32_t> d_order;
  DevBuf<int64_t> d_occ;
  DevBuf<uint32_t> d_mt, d_counter;
  DevBuf<double> d_rows, d_drift, d_msd;
  DevBuf<TrajOut> d_out;
  DevBuf<uint32_t> d_mts;
  DevBuf<int32_t> d_socc;
  DevBuf<double> d_svec, d_spsi;
  // slot pool (serving)
  bool pool = false;
  int64_t rows_capacity = 0;
  std::vector<char> slot_busy;
  std::vector<MsdTraj> msd_host;
  DevBuf<MsdTraj> d_msdtr;
  std::vector<TrajDev> trajs;
  std::vector<int64_t> msd_off;
  int64_t n_traj = 0, rows_total = 0, msd_total = 0;
  int nc_max = 0, z_max = 0;
  int nlb_max = 0;
  double msd_scale = 1.0;
  size_t smem = 0;
  int grid = 0;
  hipGraphExec_t graph = nullptr;
  float last_ms = 0.f;
  ~pycd_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    for (auto* f : fams) delete f;
  }
};

#include "batch_api.inc"

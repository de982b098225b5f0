// Batched KMC: thousands of independent Run.do_kmc_steps trajectories
// (PyCD/core.py:2668-2861) of uneven supercell size and carrier count in
// one persistent launch on gfx950.
//
// Families are cell-periodic: a supercell of nx x ny x nz unit cells with
// n_unit sites each, numbered cell-major, site = ((cx ny + cy) nz + cz)
// n_unit + sub (the order of Material.generate_sites, PyCD/core.py:154-241).
// On such a lattice the pair table, the lattice potential, the site
// energies and every hop slot are invariant under cell translations, so a
// family keeps only the table rows of the n_unit sites of cell 0 and one
// hop record per (sublattice, slot):
//   A[a, b] = C[sub(a), b shifted by -cell(a)].
// The host checks that invariance on the full table and the slot tables
// before it accepts a family.  Only the order in which a site lists its
// slots may differ from cell 0's (neighbour lists are sorted by site index,
// which wraps at the supercell boundary); one word per site holds that
// permutation (3 bits per slot), so every site keeps the slot order, and
// with it the process order of the selection, that its own table gives.
// The compact rows and the hop differences (0.6 + 2.4 MB for an 8x8x8
// corundum supercell against 302 MB for the full table) stay in the L2 of
// the XCDs the family's trajectories are dealt to, so the per-hop gathers
// are cache hits rather than HBM round trips.  Sites are carried packed:
// cx | cz << 8 | cy << 16 | sub << 24.
//
// Scheduling: one wavefront (64 lanes, one workgroup) owns one trajectory at
// a time and takes the next from a global atomic queue when it finishes, in
// the order given by the host (longest expected work first).  The queue has
// an exit condition every wave reaches (index >= n_traj), so the grid drains.
//
// Per-hop algebra is the plain path's (q.A[x,:] = phi_L[x] + psi[x]), but
// psi is maintained incrementally: when carrier m hops s -> n, every slot
// value changes by q (A[n, x] - A[s, x]), one gather per slot from the
// family's hop differences (two table gathers without them), and carrier
// m's new slots are summed from scratch (n_c gathers each); every `refresh`
// hops all of psi is recomputed from the table to bound the rounding.
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
#include "packed_sites.hpp"
#include "wave.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace pycd {

struct SubRec {       // hop slot j of sublattice s
  uint32_t tgt;       // packed cell offset (each in [0, n)) | target sub << 24
  int32_t pad0;
  double lam;         // lambda (Hartree)
  double vab;         // V_ab (Hartree)
  double tself;       // A[s, s] - A[s, n]
  double dphi;        // q_L . (A[n, :] - A[s, :])
  double dE;          // E[n] - E[s]
  double vec[3];      // hop vector (bohr)
};

struct FamDev {
  const double* C64;      // n_unit x N: table rows of the sites of cell 0
  const float* C32;
  const SubRec* subs;     // n_unit x z
  const uint32_t* perm;   // N: slot j of the site is slot
                          //    (perm >> 3 j) & 7 of its sublattice
  const double* phiL;     // n_unit
  // n_unit z x N or NULL: row s z + j holds A[n, x] - A[s, x] at the
  // column of x in s's row, n the target of slot j of sublattice s (the two
  // entries as the hop loop reads them, FP32 ones in the FP32 table mode)
  const double* D;
  int64_t N;
  int32_t nx, ny, nz, nu;
  int32_t z, pad;
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
  int32_t simd;      // SIMD key (XCC, SE, CU, SIMD) of its last wavefront
  double busy;       // wall-clock ticks the trajectory's wavefront spent on it
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
  // work queue: `order` holds n_bins bins back to back, bin b at
  // [bin_start[b], bin_start[b + 1]); counter[b] counts bin b's pops and
  // counter[n_bins] all pops.  key_bin maps a SIMD's hardware key to the bin
  // its wavefronts start from (-1: none; NULL when n_bins is 1).
  uint32_t* counter;
  const uint32_t* bin_start;
  const int16_t* key_bin;
  int32_t n_bins;
  int32_t pad1;
  // remaining hops of the wavefront in each SIMD wave slot (kSimdKeys x 16,
  // zeroed before the launch), for pacing, followed by the hand-over words
  // (see claim_mailbox); NULL: no pacing and no hand-over
  uint32_t* progress;
  // resumable state (slices): MT words, occupancy, displacement/base/drift
  uint32_t* mts;            // n_traj x 625
  uint32_t* st_occ;         // n_traj x kMaxCarriers (packed sites)
  double* st_vec;           // n_traj x 3 x (3 kMaxCarriers)
  double* st_psi;           // n_traj x kPsiStride
  int64_t slice_hops;       // <= 0: run each trajectory to its end
  int64_t refresh;          // recompute every psi from the table each
                            // `refresh` hops of a trajectory (<= 0: never)
  int32_t n_traj;
  int32_t nc_max;
  int32_t z_max;
  int32_t sub_max;          // largest n_unit z of a family the launch uses
  double nu;
};

constexpr int kMaxCarriers = 64;
constexpr int kMaxPer = 8;    // processes per lane: n_c z <= 512
constexpr int kPsiStride = kMaxCarriers * 8;  // n_c (1 + z), z <= 7
constexpr int kVecStride = 3 * 3 * kMaxCarriers;
constexpr int kMaxSubSlots = 256;  // n_unit z of a family
constexpr int kMaxCells = 255;     // cells per axis, sites per cell
constexpr double kLog2E = 1.4426950408889634;  // log2(e)

// Values every lane holds identically, moved to scalar registers.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint32_t uni(uint32_t v) {
  return uint32_t(__builtin_amdgcn_readfirstlane(int(v)));
}
__device__ inline double uni(double v) {
  const uint64_t b = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(b));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(b >> 32));
  return __longlong_as_double((int64_t(hi) << 32) | lo);
}



// The family's arrays are reached through pointers read from memory, which
// the compiler would treat as generic (flat) addresses; flat loads also
// count against the LDS wait counter, so every LDS wait would wait for the
// gathers as well.  These casts make them global loads.
template <typename T>
using gptr = const __attribute__((address_space(1))) T*;
template <typename T>
__device__ inline gptr<T> as_global(const T* p) {
  return (gptr<T>)p;
}
// (the host pass of this file sees plain pointers: its compiler has no
// constant address space to bind the records' copy constructors to)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
#else
template <typename T>
using cptr = const T*;
#endif
template <typename T>
__device__ inline cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}
template <typename T>
using gmut = __attribute__((address_space(1))) T*;
template <typename T>
__device__ inline gmut<T> as_global_mut(T* p) {
  return (gmut<T>)p;
}

__device__ inline SubRec load_rec(const FamDev& f, uint32_t k) {
  const gptr<SubRec> p = as_global(f.subs) + k;
  SubRec r;
  r.tgt = p->tgt;
  r.pad0 = 0;
  r.lam = p->lam;
  r.vab = p->vab;
  r.tself = p->tself;
  r.dphi = p->dphi;
  r.dE = p->dE;
  r.vec[0] = p->vec[0];
  r.vec[1] = p->vec[1];
  r.vec[2] = p->vec[2];
  return r;
}

template <typename TT>
__device__ inline gptr<TT> crow(const FamDev& f, uint32_t sub);
template <>
__device__ inline gptr<double> crow<double>(const FamDev& f, uint32_t sub) {
  return as_global(f.C64) + int64_t(sub) * f.N;
}
template <>
__device__ inline gptr<float> crow<float>(const FamDev& f, uint32_t sub) {
  return as_global(f.C32) + int64_t(sub) * f.N;
}

// The family's compact table (n_unit rows of N entries) reached by 32-bit
// byte offsets from one base, so that every gather takes the scalar base
// and one vector offset (the family check keeps the table under 4 GiB).
template <typename TT>
struct Tab {
  gptr<char> base;
  uint32_t row_bytes;  // N sizeof(TT)
  __device__ uint32_t row(uint32_t sub) const { return sub * row_bytes; }
  __device__ TT at(uint32_t row_off, uint32_t col) const {
    return *(gptr<TT>)(base + (row_off + col * uint32_t(sizeof(TT))));
  }
};
template <typename TT>
__device__ inline Tab<TT> tab_of(const FamDev& f) {
  Tab<TT> t;
  t.base = (gptr<char>)crow<TT>(f, 0);
  t.row_bytes = uint32_t(f.N) * uint32_t(sizeof(TT));
  return t;
}

// A[a, b] for packed sites a, b.
template <typename TT>
__device__ inline double pair_entry(const Tab<TT>& t, const GeoPk& p,
                                    uint32_t a, uint32_t b) {
  return double(t.at(t.row(a >> 24), rel_col(a, b, p)));
}

// psi[r] = sum_c q A[sites[r], occ[c]] for every slot r (lanes over slots).
template <typename TT>
__device__ inline void psi_from_table(const Tab<TT>& t, const GeoPk& gp,
                                      double q, int R, int nc,
                                      const uint32_t* sites,
                                      const uint32_t* occ, double* psi,
                                      int lane) {
  for (int r = lane; r < R; r += 64) {
    const uint32_t a = sites[r];
    const RowKey key = row_key(a, gp);
    const uint32_t row = t.row(key.sub);
    double acc = 0.0;
#pragma unroll 4
    for (int c2 = 0; c2 < nc; ++c2)
      acc = fma(q, double(t.at(row, rel_col(key, occ[c2], gp))), acc);
    psi[r] = acc;
  }
}

// Per (sublattice, slot) of the trajectory's family: the static part of a
// hop's dG0 and its rate inputs for the trajectory's charge and field, and
// the hop vector and target.  FP32 rates: 2 q (dphi + q tself) + dE, lambda,
// 1 / (4 lambda), V_ab + field shift; FP64 rates: the six terms as the
// reference adds them.
struct PTabF32 {
  double g0s;
  float lam, i4l, vs;
  uint32_t tgt;       // the hop record's packed target offset
  double vec[3];
};
struct PTabF64 {
  double dphi, tq, dE, lam, vab, shift;
  double vec[3];
  uint32_t tgt, pad;
};
template <bool kF32Rate>
using PTab = typename std::conditional<kF32Rate, PTabF32, PTabF64>::type;

// Shared-memory layout of one wavefront (sizes from the batch's maxima:
// pmax = nc_max z_max processes, rmax = nc_max (1 + z_max) slots):
//   mt     624 words (2560 B)        MT19937 block
//   psi    rmax doubles              potential of the carriers at each slot
//   ptab   sub_max PTab entries      per (sublattice, slot), see PTab
//   sites  rmax words                packed slot sites (carrier, targets)
//   occ    nc_max words              packed carrier sites
//   cperm  nc_max words              slot permutation of each carrier's site
//   pinfo  pmax words                per process: ptab index << 24 | slot
//                                    permutation of its target site
__host__ __device__ inline size_t batch_smem_bytes(int nc_max, int z_max,
                                                  bool f32rate, int sub_max) {
  const size_t rmax = size_t(nc_max) * (1 + z_max);
  const size_t pmax = size_t(nc_max) * z_max;
  const size_t pt = f32rate ? sizeof(PTabF32) : sizeof(PTabF64);
  return 2560 + 8 * rmax + pt * size_t(sub_max) + 4 * rmax +
         8 * size_t(nc_max) + 4 * pmax;
}

template <bool kF32Rate>
__device__ inline void fill_ptab(PTab<kF32Rate>& e, const SubRec& rec,
                                 double q, const double* field,
                                 bool field_on) {
  const double tq = q * rec.tself;
  const double shift =
      field_on ? 0.5 * (field[0] * rec.vec[0] + field[1] * rec.vec[1] +
                        field[2] * rec.vec[2])
               : 0.0;
  if constexpr (kF32Rate) {
    e.g0s = 2 * q * (rec.dphi + tq) + rec.dE;
    e.lam = float(rec.lam);
    e.i4l = float(0.25 / rec.lam);
    e.vs = float(rec.vab) + float(shift);
  } else {
    e.pad = 0;
    e.dphi = rec.dphi;
    e.tq = tq;
    e.dE = rec.dE;
    e.lam = rec.lam;
    e.vab = rec.vab;
    e.shift = shift;
  }
  e.vec[0] = rec.vec[0];
  e.vec[1] = rec.vec[1];
  e.vec[2] = rec.vec[2];
  e.tgt = rec.tgt;
}

// kPer: processes per lane (n_c z <= 64 kPer); kZp: slots per carrier
// (z + 1 <= kZp).  Instantiated for the batch's maxima so that the per-lane
// arrays stay small and several wavefronts fit per SIMD.
// Where this wavefront runs: XCC << 10 | SE << 7 | SH << 6 | CU << 2 | SIMD
// (13 bits), from the XCC_ID and HW_ID hardware registers.
__device__ inline uint32_t simd_key() {
  // s_getreg operand: (bits - 1) << 11 | offset << 6 | register id
  const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);   // HW_ID
  const uint32_t xcc = __builtin_amdgcn_s_getreg((15 << 11) | 20); // XCC_ID
  const uint32_t simd = (hw >> 4) & 3u, cu = (hw >> 8) & 15u,
                 sh = (hw >> 12) & 1u, se = (hw >> 13) & 7u;
  return ((xcc & 7u) << 10) | (se << 7) | (sh << 6) | (cu << 2) | simd;
}
constexpr int kSimdKeys = 1 << 13;

// simd_key() << 4 | the wavefront's slot in its SIMD (HW_ID bits 3:0).
__device__ inline uint32_t wave_slot() {
  return (simd_key() << 4) | (__builtin_amdgcn_s_getreg((3 << 11) | 4) & 15u);
}

// Wavefronts sharing a SIMD pace each other.  Each publishes the hops it
// has left and takes a priority by rank among its SIMD's wavefronts, most
// hops left highest, so that the SIMD's trajectories advance together and
// end together; left alone, the SIMD favours its older wavefronts, whose
// trajectories end first while the rest finish on a half-idle SIMD.
__device__ inline void pace(uint32_t* prog, uint32_t slot, uint32_t left,
                            int lane) {
  if (lane == 0)
    __hip_atomic_store(prog + slot, left, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t base = slot & ~15u, me = slot & 15u;
  uint32_t v = 0;
  if (lane < 16)
    v = __hip_atomic_load(prog + base + lane, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
  const bool before = lane < 16 && uint32_t(lane) != me &&
                      (v > left || (v == left && uint32_t(lane) < me));
  const int rank = __popcll(__ballot(before));
  if (rank == 0)
    __builtin_amdgcn_s_setprio(3);
  else if (rank == 1)
    __builtin_amdgcn_s_setprio(2);
  else if (rank == 2)
    __builtin_amdgcn_s_setprio(1);
  else
    __builtin_amdgcn_s_setprio(0);
}

// Index into `order` of this wavefront's next trajectory, or n_traj once
// the queue is empty (lane 0 only).  Pops from `bin` first, then moves on
// to the next bins; stops scanning as soon as every trajectory is taken.
__device__ inline uint32_t pop_ticket(const BatchArgs& a, int& bin) {
  uint32_t* taken = a.counter + a.n_bins;
  for (int k = 0; k < a.n_bins; ++k) {
    const uint32_t lo = a.bin_start[bin], n = a.bin_start[bin + 1] - lo;
    if (n > 0 &&
        __hip_atomic_load(a.counter + bin, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT) < n) {
      const uint32_t t = atomicAdd(a.counter + bin, 1u);
      if (t < n) {
        atomicAdd(taken, 1u);
        return lo + t;
      }
    }
    if (__hip_atomic_load(taken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
        uint32_t(a.n_traj))
      break;
    bin = bin + 1 == a.n_bins ? 0 : bin + 1;
  }
  return uint32_t(a.n_traj);
}

// Hand-over of running trajectories to idle wavefronts (batch launches with
// SIMD bins: the words after the pacing slots).  SIMDs of the same chip run the same
// work at speeds up to ~8% apart, so with every trajectory resident from
// the start the launch would end with the slowest SIMDs.  Instead, a
// wavefront that finds the queue empty opens a mailbox and waits; a
// wavefront whose trajectory has hops to spare saves the trajectory's
// resumable state (exact, as at the end of a slice), puts its index in a
// waiting mailbox and leaves.  Words of `hand`:
//   [0] trajectories not yet finished   [1] mailboxes taken by donors
//   [2] mailboxes opened                [3] mailboxes (0: hand-over off)
//   [kHandHdr + k] mailbox k: 0 empty, kMailClosed, or trajectory index + 1.
// Every wait ends: a mailbox is filled only while its donor's trajectory is
// unfinished, an empty mailbox is closed (by CAS) once none is, and a wait
// also gives up after kHandWaitTicks.  A trajectory is handed over at most
// once (a wavefront never passes on what it received), so at most
// grid + n_traj mailboxes are opened.
constexpr int kHandHdr = 4;
constexpr uint32_t kMailClosed = 0xffffffffu;
constexpr uint32_t kHandMinHops = 2048;          // hops left to be worth it
constexpr uint64_t kHandWaitTicks = 2000000000;  // 20 s at 100 MHz

__device__ inline uint32_t* hand_words(uint32_t* progress) {
  return progress + kSimdKeys * 16;
}

// Mailbox index of a waiting wavefront taken for a hand-over, or -1
// (lane 0 only).
__device__ inline int claim_mailbox(uint32_t* hand) {
  const uint32_t cap = hand[3];
  for (int attempt = 0; attempt < 4; ++attempt) {
    const uint32_t head = __hip_atomic_load(hand + 1, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t open = min(__hip_atomic_load(hand + 2, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT),
                              cap);
    if (head >= open) return -1;
    uint32_t expect = head;
    if (__hip_atomic_compare_exchange_strong(
            hand + 1, &expect, head + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT))
      return int(head);
  }
  return -1;
}

// Opens a mailbox and waits for a trajectory: its index, or -1 when none
// is left to hand over or hand-over is off (lane 0 only).
__device__ inline int wait_for_trajectory(uint32_t* hand) {
  const uint32_t cap = hand[3];
  if (cap == 0u) return -1;
  const uint32_t k = atomicAdd(hand + 2, 1u);
  if (k >= cap) return -1;
  uint32_t* box = hand + kHandHdr + k;
  const uint64_t t0 = wall_clock64();
  // Relaxed polls: an acquire here would invalidate the caches of the
  // wavefronts still running on every look; the caller takes one acquire
  // fence once a trajectory has arrived.
  for (uint32_t it = 0;; ++it) {
    const uint32_t v =
        __hip_atomic_load(box, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v != 0u) return int(v - 1u);
    // the shared count only every 8th look, to keep the waiters off it
    if ((it & 7u) == 7u &&
        (__hip_atomic_load(hand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             0u ||
         wall_clock64() - t0 > kHandWaitTicks)) {
      uint32_t expect = 0u;
      if (__hip_atomic_compare_exchange_strong(
              box, &expect, kMailClosed, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
              __HIP_MEMORY_SCOPE_AGENT))
        return -1;
      return int(expect - 1u);  // filled meanwhile
    }
    __builtin_amdgcn_s_sleep(127);
    __builtin_amdgcn_s_sleep(127);
  }
}

// Marks the hardware key of every SIMD that runs one of its wavefronts in
// `seen` (kSimdKeys bits).  Each wavefront sleeps `spin` x ~8k cycles
// after, so that a grid of several wavefronts per SIMD occupies all of
// them at once.
__global__ void __launch_bounds__(64) simd_probe(uint32_t* seen, int spin) {
  if (threadIdx.x == 0) {
    const uint32_t k = simd_key();
    atomicOr(seen + (k >> 5), 1u << (k & 31u));
  }
  for (int i = 0; i < spin; ++i) __builtin_amdgcn_s_sleep(127);
}

// f(std::integral_constant<int, n>) for the run-time n in [1, kMax].
template <int kMax, typename F>
__device__ __forceinline__ void with_count(int n, F&& f) {
  if constexpr (kMax >= 1) {
    if (n == kMax)
      f(std::integral_constant<int, kMax>{});
    else
      with_count<kMax - 1>(n, f);
  }
}

// Wavefronts per SIMD the register allocation targets: 4 (16 per CU, every
// trajectory of a 4096-trajectory batch resident at once, in 128 VGPRs with
// only the per-trajectory set-up spilling) for FP32 rates with at most 4
// processes per lane; 2 otherwise (the FP64-rate and large instances would
// spill heavily in 128).
#ifndef PYCD_BATCH_WAVES_PER_EU
#define PYCD_BATCH_WAVES_PER_EU(f32, per) ((f32) && (per) <= 4 ? 4 : 2)
#endif
template <typename TT, bool kF32Rate, int kPer, int kZp>
__device__ __forceinline__ void kmc_batched_body(const BatchArgs& a) {
  extern __shared__ unsigned char smem_raw[];
  const int lane = threadIdx.x;
  const int rmax = a.nc_max * (1 + a.z_max);
  uint32_t* mt = reinterpret_cast<uint32_t*>(smem_raw);
  double* psi = reinterpret_cast<double*>(smem_raw + 2560);
  PTab<kF32Rate>* ptab = reinterpret_cast<PTab<kF32Rate>*>(psi + rmax);
  uint32_t* sites = reinterpret_cast<uint32_t*>(ptab + a.sub_max);
  uint32_t* occ = sites + rmax;
  uint32_t* cperm = occ + a.nc_max;
  uint32_t* pinfo = cperm + a.nc_max;

  const uint32_t pslot = a.progress ? wave_slot() : 0u;
  int bin = 0;
  if (a.n_bins > 1) {
    const int b = a.key_bin[simd_key()];
    bin = uni(b >= 0 ? b : int(blockIdx.x % uint32_t(a.n_bins)));
  }
  int again = -1;  // a trajectory whose hand-over fell through
  for (;;) {
    int tr;
    bool stolen = false;  // handed over by another wavefront
    uint32_t ti = uint32_t(a.n_traj);
    if (again < 0) {
      if (lane == 0) ti = pop_ticket(a, bin);
      ti = uni(uint32_t(__shfl(int(ti), 0, 64)));
      bin = uni(__shfl(bin, 0, 64));
    }
    if (again >= 0) {
      tr = again;  // resumed from the state saved for the hand-over
      again = -1;
    } else if (ti < uint32_t(a.n_traj)) {
      tr = uni(a.order[ti]);
    } else {
      if (!a.progress) break;
      // wait for a hand-over only on a SIMD that has run out of work, so
      // that a trajectory moves from a busy SIMD to an idle one
      uint32_t busy_here = 0;
      if (lane < 16)
        busy_here = __hip_atomic_load(a.progress + (pslot & ~15u) + lane,
                                      __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
      if (__ballot(busy_here != 0u)) break;
      int got = -1;
      if (lane == 0) got = wait_for_trajectory(hand_words(a.progress));
      got = uni(__shfl(got, 0, 64));
      if (got < 0) break;
      // the donor's saved state, released before its mailbox write
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      tr = got;
      stolen = true;
    }
    const TrajOut o0 = a.out[tr];
    if (uni(o0.done)) {
      if (a.progress && lane == 0) atomicSub(hand_words(a.progress), 1u);
      continue;
    }
    const bool resume = uni(o0.started) != 0;
    const uint64_t w_start = wall_clock64();
    // The launch never writes the trajectory and family records: read
    // through the constant address space they are scalar loads, and their
    // fields (table base, row stride, sizes) live in scalar registers rather
    // than in vector registers read back with readfirstlane at every use.
    const TrajDev td = as_const(a.trajs)[tr];
    const FamDev f = as_const(a.fams)[uni(td.fam)];
    const Geo g{uni(f.nx), uni(f.ny), uni(f.nz), uni(f.nu)};
    const GeoPk gp = geo_pk(g);
    const Tab<TT> tab = tab_of<TT>(f);
    // the family's hop differences (absent: two table entries per slot row)
    Tab<double> dtab;
    dtab.base = (gptr<char>)as_global(f.D);
    dtab.row_bytes = uint32_t(f.N) * uint32_t(sizeof(double));
    const bool use_d = f.D != nullptr;
    const int nc = uni(td.nc), z = uni(f.z), zp = z + 1;
    const int R = nc * zp, nproc = nc * z;
    const int per = (nproc + 63) / 64;
    // this lane's processes p = lane per + i, fixed for the trajectory:
    // c | j << 6 | (slot of c's site) << 9 | (slot of the target) << 19,
    // bit 31 set for a real process
    uint32_t pdesc[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int p = lane * per + i;
      pdesc[i] = 0;
      if (i < per && p < nproc) {
        const int c = p / z, j = p - c * z;
        pdesc[i] = uint32_t(c) | (uint32_t(j) << 6) |
                   (uint32_t(c * zp) << 9) | (uint32_t(c * zp + 1 + j) << 19) |
                   0x80000000u;
      }
    }
    const double q = td.q;
    double* rows = a.rows + td.row_off;
    const int w = 3 * nc;
    const bool field_on =
        td.field[0] != 0.0 || td.field[1] != 0.0 || td.field[2] != 0.0;

    // resumable per-trajectory state in global memory (the displacement
    // and drift there are written with the state and read on resumption;
    // while the trajectory runs they are held in the carriers' lanes)
    uint32_t* mts = a.mts + int64_t(tr) * 625;
    uint32_t* socc = a.st_occ + int64_t(tr) * kMaxCarriers;
    double* svec = a.st_vec + int64_t(tr) * kVecStride;
    const gmut<double> svec_g = as_global_mut(svec);
    const gmut<double> sdisp = svec_g;
    double* sbase = svec + 3 * kMaxCarriers;
    const gmut<double> sdrift = svec_g + 6 * kMaxCarriers;
    double* spsi = a.st_psi + int64_t(tr) * kPsiStride;

    // ---- load or start the state
    const uint32_t* m0 = resume ? mts : a.mt0 + int64_t(tr) * 625;
    for (int k = lane; k < 624; k += 64) mt[k] = m0[k];
    int mti = int(m0[624]);
    for (int c = lane; c < nc; c += 64) {
      const uint32_t s =
          resume ? socc[c] : site_pack(uint32_t(a.occ0[td.occ_off + c]), g);
      occ[c] = s;
      cperm[c] = as_global(f.perm)[site_index(s, g)];
    }
    for (int k = lane; k < f.nu * z; k += 64) {
      fill_ptab<kF32Rate>(ptab[k], load_rec(f, k), q, td.field, field_on);
    }
    if (!resume)
      for (int k = lane; k < w; k += 64) {
        sbase[k] = 0.0;
        rows[k] = 0.0;  // path row 0: the initial positions
      }
    // displacement since the last path row and drift of carrier `lane`, in
    // its lane while the trajectory runs (saved with the state)
    double dsp[3] = {0.0, 0.0, 0.0}, drf[3] = {0.0, 0.0, 0.0};
    if (resume && lane < nc)
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        dsp[c3] = sdisp[3 * lane + c3];
        drf[c3] = sdrift[3 * lane + c3];
      }
    __syncthreads();
    for (int r = lane; r < R; r += 64) {
      const int c = r / zp, k = r - c * zp;
      const uint32_t s = occ[c];
      const uint32_t jj = (cperm[c] >> (3 * (k - 1))) & 7u;
      sites[r] = k == 0 ? s : site_add(s, ptab[(s >> 24) * z + jj].tgt, g);
    }
    // each process's table entry and the slot permutation of its target
    for (int p = lane; p < nproc; p += 64) {
      const int c = p / z, j = p - c * z;
      const uint32_t s = occ[c];
      const uint32_t k = (s >> 24) * z + ((cperm[c] >> (3 * j)) & 7u);
      pinfo[p] = (k << 24) |
                 as_global(f.perm)[site_index(site_add(s, ptab[k].tgt, g), g)];
    }
    __syncthreads();
    double t_clk, energy;
    // path rows are int32 (the host keeps n_path below 2^31)
    const int32_t P = int32_t(td.P);
    int32_t start;
    int64_t hops;
    int no_select;
    if (resume) {
      for (int r = lane; r < R; r += 64) psi[r] = spsi[r];
      t_clk = o0.t;
      energy = o0.energy;
      start = int32_t(o0.start);
      hops = o0.hops;
      no_select = o0.no_select;
    } else {
      psi_from_table<TT>(tab, gp, q, R, nc, sites, occ, psi, lane);
      // initial energy: ewald_neut + q_L A q_L + sum_c q (2 phi_L + psi_self)
      double e_loc = 0.0;
      for (int c = lane; c < nc; c += 64) {
        const uint32_t s = occ[c];
        double inner = 0.0;
        for (int c2 = 0; c2 < nc; ++c2)
          inner = fma(q, pair_entry<TT>(tab, gp, s, occ[c2]), inner);
        e_loc += q * (2.0 * as_global(f.phiL)[s >> 24] + inner);
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
    // next hop count that is a positive multiple of `refresh`
    int64_t next_refresh = INT64_MAX;
    if (a.refresh > 0)
      next_refresh =
          hops == 0 ? a.refresh : (hops + a.refresh - 1) / a.refresh * a.refresh;
    // FP32 rates: nu exp(-gs / kT) = 2^(gs (-log2(e) / kT) + log2(nu)),
    // one FMA ahead of v_exp_f32
    const float exp_scale = float(-kLog2E / td.kT);
    const float exp_bias = float(log2(a.nu));
    const double inv_dt = 1.0 / td.dt;

    // Uniforms of the coming hops, one hop per lane: at a refill (mti = m)
    // lane l takes u1 and log(u2) (-inf for u2 = 0) of the hop that draws
    // words m + 4 l .. m + 4 l + 3 of the MT block in LDS, and lane c_next
    // holds the next hop's.  Only hops whose four words all lie in the
    // block are cached, so the block is never regenerated ahead of the hops
    // that consume it; a hop whose words straddle the block's end draws
    // them directly.  Each lane takes one FP64 log per up to 64 hops
    // instead of every lane taking one per hop.
    double c_u1 = 0.0, c_lu2 = 0.0;
    int c_next = 0, c_n = 0;
    // outputs and resumable state: at the trajectory's end, or at a hop
    // boundary when it is handed over
    auto save = [&](bool done_now) {
      // every carrier slot of the output, zero past the trajectory's own
      static_assert(kMaxCarriers == 64, "one carrier per lane");
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        a.drift[int64_t(tr) * 3 * kMaxCarriers + 3 * lane + c3] =
            lane < nc ? drf[c3] : 0.0;
        if (lane < nc) {
          sdisp[3 * lane + c3] = dsp[c3];
          sdrift[3 * lane + c3] = drf[c3];
        }
      }
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
        o.done = done_now ? 1 : 0;
        o.started = 1;
        o.simd = int32_t(simd_key());
        o.busy = (resume ? o0.busy : 0.0) + double(wall_clock64() - w_start);
        a.out[tr] = o;
      }
      __syncthreads();
    };
    int box = -1;  // mailbox of the wavefront taking the trajectory over
    // Hops this launch may still make: to max_hops (which finishes the
    // trajectory) and to the end of the slice (which pauses it).  The hop
    // loop counts them down in 32 bits and asks again when they run out, so
    // a budget above 2^31 - 1 is taken in parts.
    auto budget = [&]() {
      int64_t h = INT64_MAX;
      if (td.max_hops > 0) h = td.max_hops - hops;
      if (a.slice_hops > 0) h = min(h, a.slice_hops - slice_done);
      return h;
    };
    uint32_t left = 0;
    for (;;) {
      // Conditions on the lane and on the trajectory's sizes are formed
      // again each hop from these copies, which the compiler cannot move
      // out of the loop.  Hoisted, each such condition is a 64-bit lane mask
      // held across the loop, and there are more of them than registers.
      int lane_h = lane, per_h = per, z_h = z, nc_h = nc, R_h = R;
      asm volatile("" : "+v"(lane_h));
      asm volatile("" : "+s"(per_h), "+s"(z_h), "+s"(nc_h), "+s"(R_h));
      uint32_t pd[kPer];
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        pd[i] = pdesc[i];
        asm volatile("" : "+v"(pd[i]));
      }
      if (left == 0) {
        asm volatile("");  // the 64-bit budget only when the count runs out
        const int64_t b = budget();
        if (b <= 0) {
          finished = td.max_hops > 0 && hops >= td.max_hops;
          break;
        }
        left = uint32_t(min<int64_t>(b, INT32_MAX));
      }
      --left;
      ++slice_done;
      if (a.progress && (uint32_t(hops) & 63u) == 0) {
        pace(a.progress, pslot, left + 1, lane_h);
        if (!stolen && left >= kHandMinHops) {
          if (lane_h == 0) box = claim_mailbox(hand_words(a.progress));
          box = uni(__shfl(box, 0, 64));
          if (box >= 0) break;  // hand over at this hop boundary
        }
      }
      if (hops == next_refresh) {
        next_refresh += a.refresh;
        // bound the rounding carried by the incremental updates
        psi_from_table<TT>(tab, gp, q, R_h, nc_h, sites, occ, psi, lane_h);
        __syncthreads();
      }
      // ---- two uniforms, as random.random() draws them
      if (c_next >= c_n && mti <= kMtN - 4) {
        c_n = min(64, (kMtN - mti) / 4);
        c_next = 0;
        if (lane_h < c_n) {
          const int w = mti + 4 * lane_h;
          c_u1 = mt_uniform(mt_temper(mt[w]), mt_temper(mt[w + 1]));
          const double u2 = mt_uniform(mt_temper(mt[w + 2]), mt_temper(mt[w + 3]));
          c_lu2 = u2 > 0.0 ? log(u2) : -INFINITY;
        }
      }
      double u1, lu2;
      if (c_next < c_n) {
        u1 = read_lane(c_u1, c_next);
        lu2 = read_lane(c_lu2, c_next);
        ++c_next;
        mti += 4;
      } else {
        uint32_t wd[4];
        mt_next4(mt, mti, lane_h, wd);
        u1 = mt_uniform(wd[0], wd[1]);
        const double u2 = mt_uniform(wd[2], wd[3]);
        lu2 = u2 > 0.0 ? log(u2) : -INFINITY;
      }
      // ---- rates from shared memory only
      double kv[kPer], gv[kPer];
      double ksum = 0.0;
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        kv[i] = 0.0;
        gv[i] = 0.0;
      }
      // one straight-line body per processes-per-lane count, so that the
      // processes' LDS reads are in flight together
      with_count<kPer>(per_h, [&](auto n_per) {
        constexpr int kP = decltype(n_per)::value;
#pragma unroll
        for (int i = 0; i < kP; ++i) {
          const uint32_t d = pd[i];
          const bool real = d >> 31;
          const int p = real ? lane_h * kP + i : 0;
          const PTab<kF32Rate>& e = ptab[pinfo[p] >> 24];
          const double dpsi =
              psi[(d >> 19) & 0x3ffu] - psi[(d >> 9) & 0x3ffu];
          double g0;
          if constexpr (kF32Rate) {
            g0 = e.g0s + 2 * q * dpsi;
            const float gf = float(g0);
            const float gs = (e.lam + gf) * (e.lam + gf) * e.i4l - e.vs;
            const float kf =
                __builtin_amdgcn_exp2f(fmaf(gs, exp_scale, exp_bias));
            kv[i] = double(real ? kf : 0.0f);
          } else {
            g0 = 2 * q * ((e.dphi + dpsi) + e.tq) + e.dE;
            const double lam = e.lam;
            const double gs =
                (((lam + g0) * (lam + g0) / (4 * lam)) - e.vab) - e.shift;
            const double k = a.nu * numpy_e_pow(-gs / td.kT);
            kv[i] = real ? k : 0.0;
          }
          gv[i] = g0;
          ksum = i == 0 ? kv[0] : ksum + kv[i];
        }
      });
      // ---- FP64 prefix and selection.  FP64 rates: the prefix of the
      // quotients k / k_tot against u1, as the reference forms it; FP32
      // rates: one scan of the rates themselves against u1 k_tot.
      double k_tot, thr, cum;
      double xq[kPer];
      if (kF32Rate) {
        const double incl = wave_incl_scan(ksum);
        k_tot = read_lane(incl, 63);
        cum = dpp_f64<kDppWaveShr1, 0xf, true>(incl);
        thr = u1 * k_tot;
#pragma unroll
        for (int i = 0; i < kPer; ++i) xq[i] = kv[i];
      } else {
        k_tot = wave_sum(ksum);
        double xs = 0.0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
          xq[i] = kv[i] / k_tot;
          xs += xq[i];
        }
        cum = wave_excl_scan(xs);
        thr = u1;
      }
      // The selected process is the first (lane, i), lane-major, whose
      // running sum cum + xq[0] + ... + xq[i] exceeds thr.  The terms are
      // non-negative, so each lane's running sums do not decrease: the
      // count of them that stay <= thr is the index of the lane's first
      // crossing (a lane crosses iff the count is below per, and its first
      // process is real: processes past nproc add zero at the lane's end).
      int nb = 0;
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        if (i < per_h) {
          cum += xq[i];
          nb += cum > thr ? 0 : 1;
        }
      }
      const unsigned long long mask =
          __ballot((pd[0] >> 31) != 0u && nb < per_h);
      int src, isel;
      if (mask) {
        src = uni(__ffsll(static_cast<long long>(mask)) - 1);
        isel = __builtin_amdgcn_readlane(nb, src);
      } else {
        // rounding left every running sum <= thr: take the last process
        src = (nproc - 1) / per_h;
        isel = nproc - 1 - src * per_h;
        ++no_select;
      }
      const int sel = src * per_h + isel;
      uint32_t dsel = 0;
      double gsel = 0.0, ksel = 0.0;
#pragma unroll
      for (int i = 0; i < kPer; ++i)
        if (i == isel) {
          dsel = uint32_t(__builtin_amdgcn_readlane(int(pd[i]), src));
          gsel = read_lane(gv[i], src);
          ksel = read_lane(kv[i], src);
        }
      // ---- clock
      int32_t end;
      if (lu2 > -INFINITY) {  // u2 > 0
        t_clk -= lu2 / k_tot;
        // path step int(t / dt) of the correctly rounded quotient: below
        // 1e6, t (1/dt) is within 4e-10 of it, so more than 1e-9 away from
        // an integer its integer part is the quotient's; otherwise divide
        const double qf = t_clk * inv_dt;
        const double fr = qf - floor(qf);
        if (fr > 1e-9 && fr < 1.0 - 1e-9 && qf < 1.0e6) {
          end = int32_t(qf);
        } else {
          asm volatile("");  // the division only where it is needed
          const double qq = t_clk / td.dt;
          end = qq < double(P) ? int32_t(qq) : P;
        }
      } else {
        t_clk = INFINITY;
        end = P;
      }
      energy += gsel;
      // ---- one round of independent loads for the whole update
      const int m = int(dsel & 63u);
      const int j = int((dsel >> 6) & 7u);
      const uint32_t s_old = uni(sites[m * zp]);
      const uint32_t s_new = uni(sites[m * zp + 1 + j]);
      const uint32_t sub_new = s_new >> 24;
      const uint32_t info_sel = uni(pinfo[sel]);
      const uint32_t perm_new = info_sel & 0xffffffu;
      const uint32_t rec_old = info_sel >> 24;
      // table entries of the new site's processes and the slot
      // permutations of their targets (lanes < z), issued first
      uint32_t info_new = 0, nsite = 0;
      if (lane_h < z_h) {
        const uint32_t k = sub_new * z_h + ((perm_new >> (3 * lane_h)) & 7u);
        nsite = site_add(s_new, ptab[k].tgt, g);
        info_new = (k << 24) | as_global(f.perm)[site_index(nsite, g)];
      }
      uint32_t nsel[kZp - 1];  // targets of the new site, in its slot order
#pragma unroll
      for (int k = 0; k < kZp - 1; ++k)
        nsel[k] = k < z_h ? uint32_t(__builtin_amdgcn_readlane(int(nsite), k))
                          : 0u;
      const uint32_t rn = tab.row(sub_new), ro = tab.row(s_old >> 24);
      const RowKey key_new = row_key(s_new, gp), key_old = row_key(s_old, gp);
      const uint32_t rd = dtab.row(rec_old);
      constexpr int kRPer = kMaxCarriers * kZp / 64;
      // Every slot's change q (A[n, x] - A[s, x]) is loaded before any is
      // used, so that the slots' gathers are in flight together: one entry
      // of the family's hop differences (row rec_old, x's column in s_old's
      // row) where it has them, else both table entries (an - double(ao),
      // ao = 0 with differences)
      double an[kRPer];
      TT ao[kRPer];
      // the slot sites of this lane's rows, read together (one LDS wait,
      // not one per row)
      uint32_t xs[kRPer];
#pragma unroll
      for (int i = 0; i < kRPer; ++i) {
        const int r = lane_h + 64 * i;
        xs[i] = sites[r < R_h ? r : 0];
      }
#pragma unroll
      for (int i = 0; i < kRPer; ++i) {
        const int r = lane_h + 64 * i;
        // set and read only for r < R (no zero fill of unused rows)
        if (r < R_h) {
          if (use_d) {
            an[i] = dtab.at(rd, rel_col(key_old, xs[i], gp));
            ao[i] = TT(0);
          } else {
            an[i] = double(tab.at(rn, rel_col(key_new, xs[i], gp)));
            ao[i] = tab.at(ro, rel_col(key_old, xs[i], gp));
          }
        }
      }
      // fresh sums of the new slots (not the one that returns to s_old):
      // one carrier per lane (n_c <= 64)
      // (entries loaded here, used after the apply: the loads overlap)
      constexpr int kFresh = (kZp - 1 + 3) / 4 * 4;
      TT fent[kFresh];
      {
        const bool has = lane_h < nc_h;
        const uint32_t o = has ? (lane_h == m ? s_new : occ[lane_h]) : 0u;
#pragma unroll
        for (int k = 0; k < kFresh; ++k) {
          fent[k] = TT(0);
          if (k < kZp - 1 && k < z_h && has && nsel[k] != s_old)
            fent[k] = tab.at(tab.row(nsel[k] >> 24), rel_col(nsel[k], o, gp));
        }
      }
      // hop vector of carrier m (the same address in every lane)
      double hv[3];
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) hv[c3] = ptab[rec_old].vec[c3];
      // ---- apply (process m's old entries are dead once it is selected)
      if (lane_h < z_h) pinfo[m * z_h + lane_h] = info_new;
#pragma unroll
      for (int i = 0; i < kRPer; ++i) {
        const int r = lane_h + 64 * i;
        if (r < R_h) psi[r] += q * (an[i] - double(ao[i]));
      }
      if (lane_h == m) {
#pragma clang fp contract(off)
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          dsp[c3] += hv[c3];
          if (field_on) drf[c3] = drf[c3] + hv[c3] * ksel;
        }
      }
      __syncthreads();
      const double v_new = uni(psi[m * zp + 1 + j]);
      const double v_old = uni(psi[m * zp]);
      double part[kFresh], fresh_v[kFresh];
#pragma unroll
      for (int k = 0; k < kFresh; ++k) part[k] = q * double(fent[k]);
#pragma unroll
      for (int k = 0; k < kFresh; k += 4) wave_sum4(part + k, fresh_v + k);
#pragma unroll
      for (int k = 0; k < kZp - 1; ++k)
        fresh_v[k] = (k < z_h && nsel[k] != s_old) ? fresh_v[k] : v_old;
      __syncthreads();
      if (lane_h == 0) {
        occ[m] = s_new;
        cperm[m] = perm_new;
        sites[m * zp] = s_new;
        psi[m * zp] = v_new;
#pragma unroll
        for (int k = 0; k < kZp - 1; ++k)
          if (k < z_h) {
            sites[m * zp + 1 + k] = nsel[k];
            psi[m * zp + 1 + k] = fresh_v[k];
          }
      }
      // ---- path rows [start, end), written by the owner lane of each entry
      ++hops;
      bool full = false;
      if (end >= start + 1) {
        if (end >= P) end = P;
        if (lane_h < nc_h)
#pragma unroll
          for (int c3 = 0; c3 < 3; ++c3) {
            const int k = 3 * lane_h + c3;
            const double v = sbase[k] + dsp[c3];
            dsp[c3] = 0.0;
            for (int32_t r = start; r < end; ++r)
              rows[int64_t(r) * w + k] = v;
            sbase[k] = v;
          }
        start = end;
        full = end >= P;
      }
      __syncthreads();
      if (full) {
        finished = true;
        break;
      }
    }
    if (a.progress) {  // nothing left: out of its SIMD mates' way
      if (lane == 0)
        __hip_atomic_store(a.progress + pslot, 0u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_setprio(0);
    }
    save(box < 0 && finished);
    if (box >= 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      int handed = 0;
      if (lane == 0) {
        uint32_t expect = 0u;
        handed = __hip_atomic_compare_exchange_strong(
            hand_words(a.progress) + kHandHdr + box, &expect,
            uint32_t(tr) + 1u, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT);
      }
      // the trajectory lives on in the wavefront it was handed to
      if (uni(__shfl(handed, 0, 64))) return;
      again = tr;  // that wavefront stopped waiting: resume it here
      continue;
    }
    if (a.progress && lane == 0) atomicSub(hand_words(a.progress), 1u);
  }
}

template <typename TT, bool kF32Rate, int kPer, int kZp>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(PYCD_BATCH_WAVES_PER_EU(kF32Rate, kPer))))
kmc_batched_kernel(BatchArgs a) {
  kmc_batched_body<TT, kF32Rate, kPer, kZp>(a);
}

// The serving pool's slices: the same body under its own name, so that a
// profile tells the batch launch from the pool's launches.
template <typename TT, bool kF32Rate, int kPer, int kZp>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(PYCD_BATCH_WAVES_PER_EU(kF32Rate, kPer))))
kmc_pool_kernel(BatchArgs a) {
  kmc_batched_body<TT, kF32Rate, kPer, kZp>(a);
}

// Squared displacement per trajectory, averaged over its carriers:
//   msd[t, tau] = mean_c mean_{t0 < P_t - tau} |r_c(t0 + tau) - r_c(t0)|^2
// (P_t: the path rows the run wrote).  Two kernels.
//
// msd_unit_kernel: one workgroup of kMsdWaves wavefronts per (trajectory,
// carrier) unit.  The carrier's positions pass through LDS (x, y and z
// planes) in windows read from global memory once each.  A pass takes
// kMsdPass lags in blocks of kMsdBlock; in block b lane l owns the kMsdL
// consecutive lags pass0 + kMsdBlock b + kMsdL l + i.  For a block, a
// wavefront walks consecutive origins t0 and keeps r(t0 + lag) of its lane's
// lags in a ring of kMsdRing registers, so a step reads the origin (one
// address for the whole wavefront) and one new position, and forms kMsdL
// squared distances.  Origin t0 has work in block b only while
// t0 < P - pass0 - kMsdBlock b, so the blocks of high lags have fewer
// origins; the segment's (block, origin) steps are cut into kMsdWaves equal
// runs, one per wavefront (a run may end one block and start the next), and
// the steps near the end of the written rows, where some of a lane's lags
// fall past them, mask those lags.
// Rows are held with one pad double after every 4, so the rows of lanes l
// and l + 1 (4 apart) sit 5 doubles apart and the lanes of an LDS access
// fall on distinct banks.  Runs start at multiples of 4 rows, so within a
// run every LDS address is a per-lane (or per-wavefront) base plus an
// immediate offset.  For the first pass the window holds rows
// [T0, T0 + S + span - 1), which contain the segment's S origins; later
// passes keep the origins in their own S rows ahead of the window.  Sums
// over origins, unscaled, per (unit, lag) to scratch: each wavefront's sums
// added in wavefront order.
//
// msd_reduce_kernel: one thread per (trajectory, lag) sums the carriers in
// order and scales by scale^2 / (P_t - tau) / n_c: the same order in every
// launch, so results repeat bit for bit.
//
// Workgroup b runs on XCD b mod 8 (workgroups are dealt to the XCDs in
// turn); the host deals each trajectory's units to one XCD, so the cache
// lines of path rows that its carriers share come from HBM once.
struct MsdTraj {
  int64_t row_off;
  int64_t P;
  int32_t nc;
  int32_t traj;     // index into the trajectory results (rows written)
  int64_t out_off;
  int64_t M;
  int64_t sbase;    // scratch offset of carrier 0's M lag sums
};

constexpr int kMsdL = 4;                       // lags per lane and block
constexpr int kMsdBlock = 64 * kMsdL;          // lags per block
constexpr int kMsdBlocks = 2;                  // blocks per pass
constexpr int kMsdPass = kMsdBlocks * kMsdBlock;  // lags per pass
constexpr int kMsdRing = 8;                    // ring registers (> kMsdL)
constexpr int kMsdWin = 1024;                  // path rows held in LDS
constexpr int kMsdPlane = kMsdWin + kMsdWin / 4;  // a plane with its pads
// LDS position of window row r
__device__ inline int msd_row(int r) { return r + (r >> 2); }
constexpr int kMsdXcds = 8;
constexpr int kMsdWaves = 4;           // wavefronts per unit
constexpr int kMsdThreads = 64 * kMsdWaves;
static_assert(kMsdRing % 4 == 0, "a ring turn keeps the runs' row phase");

// One wavefront's run over the relative origins [ra, rb) of a block (ra a
// multiple of 4).  L: the window, window row r at L[3 msd_row(r) + k] for
// coordinate k; lo: the lane's first lag relative to the pass; hi: the
// block's last lag below the pass's span; lim: P - pass0 - T0, the relative
// origin bound; base: window row of relative origin 0's lag-0 row.  It reads
// the rows below base + rb + lo + kMsdRing and the origins below rb + 2.
__device__ inline void msd_run(const double* L, int ra, int rb, int lo,
                               int hi, int lim, int base, double* acc) {
  constexpr int kR = kMsdRing;
  // rows are read at vb + immediate and origins at ob + immediate, both at
  // multiples of 4 rows: msd_row(4a + p) = 5a + p + (p >> 2)
  int vb = 3 * (((base + ra + lo) >> 2) * 5);
  int ob = 3 * ((ra >> 2) * 5);
  auto off = [](int p) { return 3 * (p + (p >> 2)); };
  double wx[kR], wy[kR], wz[kR], ox[2], oy[2], oz[2];
#pragma unroll
  for (int p = 0; p < kR; ++p) {
    wx[p] = L[vb + off(p)];
    wy[p] = L[vb + off(p) + 1];
    wz[p] = L[vb + off(p) + 2];
  }
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    ox[p] = L[ob + off(p)];
    oy[p] = L[ob + off(p) + 1];
    oz[p] = L[ob + off(p) + 2];
  }
  // below fast_end every lane's lags up to hi have rows inside the written
  // rows
  const int fast_end = min(rb, lim - hi);
  int o = ra;
  // at step j of a turn slot (p % kR) holds r(o + lo + p) for
  // p = j .. j + kR - 1 and origin slot (j % 2) holds r(o + j); after its
  // pairs the step loads position j + kR and origin j + 2 into the slots
  // that position j and origin j left (the scheduling barrier keeps those
  // loads after the pairs, so a slot keeps its registers)
  for (; o + kR <= fast_end; o += kR) {
#pragma unroll
    for (int j = 0; j < kR; ++j) {
#pragma unroll
      for (int i = 0; i < kMsdL; ++i) {
        const int q = (j + i) % kR;
        const double dx = wx[q] - ox[j & 1], dy = wy[q] - oy[j & 1],
                     dz = wz[q] - oz[j & 1];
        acc[i] = fma(dz, dz, fma(dy, dy, fma(dx, dx, acc[i])));
      }
      __builtin_amdgcn_sched_barrier(0);
      wx[j] = L[vb + off(j + kR)];
      wy[j] = L[vb + off(j + kR) + 1];
      wz[j] = L[vb + off(j + kR) + 2];
      ox[j & 1] = L[ob + off(j + 2)];
      oy[j & 1] = L[ob + off(j + 2) + 1];
      oz[j & 1] = L[ob + off(j + 2) + 2];
    }
    vb += 3 * (kR / 4 * 5);
    ob += 3 * (kR / 4 * 5);
  }
  // masked steps: pair i of step j counts while lo + i < lim - o - j, for
  // each pair a first part of the lanes (an exec mask, not a select)
  auto masked_step = [&](int j) {
    const int c = lim - (o + j);
#pragma unroll
    for (int i = 0; i < kMsdL; ++i) {
      const int q = (j + i) % kR;
      const double dx = wx[q] - ox[j & 1], dy = wy[q] - oy[j & 1],
                   dz = wz[q] - oz[j & 1];
      if (lo + i < c)
        acc[i] = fma(dz, dz, fma(dy, dy, fma(dx, dx, acc[i])));
    }
    __builtin_amdgcn_sched_barrier(0);
    wx[j] = L[vb + off(j + kR)];
    wy[j] = L[vb + off(j + kR) + 1];
    wz[j] = L[vb + off(j + kR) + 2];
    ox[j & 1] = L[ob + off(j + 2)];
    oy[j & 1] = L[ob + off(j + 2) + 1];
    oz[j & 1] = L[ob + off(j + 2) + 2];
  };
  // whole turns, then the last part of a turn
  for (; o + kR <= rb; o += kR) {
#pragma unroll
    for (int j = 0; j < kR; ++j) masked_step(j);
    vb += 3 * (kR / 4 * 5);
    ob += 3 * (kR / 4 * 5);
  }
#pragma unroll
  for (int j = 0; j < kR; ++j) {
    if (o + j >= rb) break;
    masked_step(j);
  }
}

__global__ void __launch_bounds__(kMsdThreads)
__attribute__((amdgpu_waves_per_eu(4, 4)))
msd_unit_kernel(const double* __restrict__ rows,
                const MsdTraj* __restrict__ tr,
                const TrajOut* __restrict__ outs,
                const int32_t* __restrict__ units,
                double* __restrict__ scratch) {
  // the window (x, y, z of a row side by side); after a pass its space
  // takes the wavefronts' lag sums
  __shared__ double L[3 * kMsdPlane];
  static_assert(kMsdWaves * kMsdPass <= 3 * kMsdPlane, "lag sums fit");
  const int32_t u = units[blockIdx.x];
  if (u < 0) return;  // padding of an XCD's list (the whole workgroup)
  const MsdTraj d = tr[u >> 6];
  const int c = u & 63;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  // only the rows written by the run: [0, start) (all P when the path filled)
  const int64_t written = outs[d.traj].start;
  const int64_t P = written < 1 ? 1 : (written < d.P ? written : d.P);
  const int64_t w3 = 3 * int64_t(d.nc);
  const gptr<double> src = (gptr<double>)(rows + d.row_off + 3 * c);
  double* out = scratch + d.sbase + int64_t(c) * d.M;
  for (int64_t pass0 = 0; pass0 < d.M; pass0 += kMsdPass) {
    const int span = int(min<int64_t>(kMsdPass, d.M - pass0));
    const int nblk = (span + kMsdBlock - 1) / kMsdBlock;
    double acc[kMsdBlocks][kMsdL];
#pragma unroll
    for (int b = 0; b < kMsdBlocks; ++b)
#pragma unroll
      for (int i = 0; i < kMsdL; ++i) acc[b][i] = 0.0;
    // origins per segment (a multiple of 4), and where the window rows sit
    // in LDS: a run reads rows below wofs + S + kMsdBlock nblk + 4, inside
    // the window
    const bool first = pass0 == 0;
    const int room = kMsdWin - kMsdBlock * nblk - 4;
    const int S = (first ? room : room / 2) & ~3;
    const int wofs = first ? 0 : S;  // window row of path row T0 + pass0
    // origins t0 < P - pass0 have a lag of this pass inside the rows
    const int64_t t_end = P - pass0;
    for (int64_t T0 = 0; T0 < t_end; T0 += S) {
      const int64_t T1 = min<int64_t>(T0 + S, t_end);
      // ---- stage the window (and, after the first pass, the origins)
      const int64_t wlo = T0 + pass0;
      const int nw = int(min<int64_t>(P, T1 - 1 + pass0 + span) - wlo);
      __syncthreads();
      for (int r = tid; r < nw; r += kMsdThreads) {
        const int64_t o = (wlo + r) * w3;
        const int pr = 3 * msd_row(wofs + r);
        L[pr] = src[o];
        L[pr + 1] = src[o + 1];
        L[pr + 2] = src[o + 2];
      }
      if (!first)
        for (int r = tid; r < T1 - T0; r += kMsdThreads) {
          const int64_t o = (T0 + r) * w3;
          const int pr = 3 * msd_row(r);
          L[pr] = src[o];
          L[pr + 1] = src[o + 1];
          L[pr + 2] = src[o + 2];
        }
      __syncthreads();
      // block b has the relative origins [0, n_b); the blocks' steps, each
      // count rounded up to a multiple of 4, are cut into kMsdWaves runs
      const int lim = int(min<int64_t>(t_end - T0, int64_t(1) << 30));
      const int seg = int(T1 - T0);
      int n[kMsdBlocks], tot = 0;
#pragma unroll
      for (int b = 0; b < kMsdBlocks; ++b) {
        n[b] = b < nblk ? max(0, min(seg, lim - kMsdBlock * b)) : 0;
        tot += (n[b] + 3) & ~3;
      }
      const int Q = ((tot + kMsdWaves - 1) / kMsdWaves + 3) & ~3;
      const int a = wave * Q, e = min(a + Q, tot);
      int boff = 0;
#pragma unroll
      for (int b = 0; b < kMsdBlocks; ++b) {
        const int ra = max(a - boff, 0), rb = min(e - boff, n[b]);
        if (ra < rb)
          msd_run(L, ra, rb, kMsdBlock * b + kMsdL * lane,
                  min(span, kMsdBlock * (b + 1)) - 1, lim, wofs, acc[b]);
        boff += (n[b] + 3) & ~3;
      }
    }
    // the wavefronts' sums of each lag added in wavefront order
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMsdBlocks; ++b)
#pragma unroll
      for (int i = 0; i < kMsdL; ++i)
        L[wave * kMsdPass + kMsdBlock * b + kMsdL * lane + i] = acc[b][i];
    __syncthreads();
    for (int l = tid; l < span; l += kMsdThreads) {
      double v = L[l];
#pragma unroll
      for (int w = 1; w < kMsdWaves; ++w) v += L[w * kMsdPass + l];
      out[pass0 + l] = v;
    }
  }
}

__global__ void __launch_bounds__(256)
msd_reduce_kernel(const MsdTraj* __restrict__ tr, int64_t n_traj,
                  const TrajOut* __restrict__ outs,
                  const double* __restrict__ scratch, double scale,
                  double* __restrict__ out) {
  const int64_t k = blockIdx.x;
  if (k >= n_traj) return;
  const MsdTraj d = tr[k];
  const int64_t written = outs[d.traj].start;
  const int64_t P = written < 1 ? 1 : (written < d.P ? written : d.P);
  const double s2 = scale * scale;
  for (int64_t tau = threadIdx.x; tau < d.M; tau += blockDim.x) {
    double v = 0.0;
    if (tau >= 1 && tau < P) {
      double acc = 0.0;
      for (int c = 0; c < d.nc; ++c) acc += scratch[d.sbase + c * d.M + tau];
      v = acc * s2 / double(P - tau) / double(d.nc);
    }
    out[d.out_off + tau] = v;
  }
}

__global__ void fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Hand-over words before a launch: `active` unfinished trajectories, no
// mailbox taken or opened, `cap` mailboxes (0: off), every mailbox empty.
__global__ void hand_reset(uint32_t* h, int64_t n, uint32_t active,
                           uint32_t cap) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    h[i] = i == 0 ? active : (i == 3 ? cap : 0u);
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
__global__ void to_f32(const double* __restrict__ a, float* __restrict__ b,
                       int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = float(a[i]);
}

// Largest |A[i, j] - C[sub(i), rel(i, j)]| over the full table A, whose
// first n_unit rows are C, as the bit pattern of a non-negative double
// (ordered like the values).  One workgroup of 256 per 4 rows.
__global__ void __launch_bounds__(256)
compact_check(const double* __restrict__ A, int64_t N, Geo g,
              unsigned long long* __restrict__ worst) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double m = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * 4 + wave; i < N;
       i += int64_t(gridDim.x) * 4) {
    const uint32_t a = site_pack(uint32_t(i), g);
    const double* row = A + i * N;
    const double* crow0 = A + int64_t(a >> 24) * N;
    for (int64_t j = lane; j < N; j += 64) {
      const uint32_t b = site_pack(uint32_t(j), g);
      m = nan_max(m, fabs(row[j] - crow0[rel_col(a, b, geo_pk(g))]));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    m = nan_max(m, __shfl_xor(m, off, 64));
  if (lane == 0)
    atomicMax(worst, static_cast<unsigned long long>(__double_as_longlong(m)));
}

}  // namespace pycd

// ------------------------------------------------------------------- host --
using namespace pycd;

struct BatchFamily {
  DevBuf<double> C64, phiL, D;
  DevBuf<float> C32;
  DevBuf<SubRec> subs;
  DevBuf<uint32_t> perm;
  int64_t N = 0;
  int z = 0, mz = 0;
  Geo g{1, 1, 1, 1};
  double qLAqL = 0.0;
  std::vector<char> hop_sub;  // per sublattice: its sites hop
  size_t device_bytes = 0;
};

struct pycd_batch {
  pycd_ctx* ctx = nullptr;
  std::vector<BatchFamily*> fams;
  DevBuf<FamDev> d_fams;
  int table_prec = 0;  // 0 FP64, 1 FP32
  bool use_diff = true;  // launches read the families' hop differences
  int64_t refresh = 1024;   // set from the rate precision at creation
  int rate_prec = 0;   // 0 FP64, 1 FP32
  double nu = 0.0;
  // per-run buffers
  DevBuf<TrajDev> d_trajs;
  DevBuf<int32_t> d_order;
  DevBuf<int64_t> d_occ;
  DevBuf<uint32_t> d_mt, d_counter, d_bins;
  DevBuf<int16_t> d_keybin;
  // pacing slots, one per SIMD wave slot, then the hand-over words (binned
  // runs)
  DevBuf<uint32_t> d_prog;
  int n_bins = 1;
  uint32_t hand_cap = 0;  // mailboxes (0 for the serving pool)
  bool handover = true;
  DevBuf<double> d_rows, d_drift, d_msd;
  DevBuf<TrajOut> d_out;
  DevBuf<uint32_t> d_mts;
  DevBuf<uint32_t> d_socc;
  DevBuf<double> d_svec, d_spsi;
  // slot pool (serving)
  bool pool = false;
  int64_t rows_capacity = 0;
  std::vector<char> slot_busy;
  std::vector<MsdTraj> msd_host;
  DevBuf<MsdTraj> d_msdtr;
  // MSD launch plan: (trajectory, carrier) units in workgroup order and
  // the per-unit lag sums
  DevBuf<int32_t> d_msd_units;
  DevBuf<double> d_msd_scratch;
  int64_t msd_units = 0;
  std::vector<TrajDev> trajs;
  std::vector<int64_t> msd_off;
  int64_t n_traj = 0, rows_total = 0, msd_total = 0;
  int nc_max = 0, z_max = 0;
  int sub_max = kMaxSubSlots;
  double msd_scale = 1.0;
  size_t smem = 0;
  int grid = 0;
  hipGraphExec_t graph = nullptr;
  float last_ms = 0.f;
  float kernel_ms = 0.f;  // the KMC kernel alone, last eager run or step
  ~pycd_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    for (auto* f : fams) delete f;
  }
};

#include "batch_api.inc"


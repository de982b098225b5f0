/*
 * pycd_hip.h -- C ABI of libpycd_hip.so, the MI355X (gfx950) implementation
 * of PyCD's kinetic Monte Carlo hot path.
 *
 * PyCD itself is pure Python and has no FFI: its hot path is three methods
 * of its Python classes.  Each entry point below replaces the numerical core
 * of one of them and is bound from Python with ctypes (pycd_amd/_native.py):
 *
 *   pycd_pair_table       <- System.get_precomputed_array
 *                            (PyCD/core.py:1611-1679; kernels for
 *                            pot_r_ewald :799-814, pot_k_ewald :853-878)
 *   pycd_kmc_*            <- Run.do_kmc_steps loop body
 *                            (PyCD/core.py:2787-2861 with get_process_rates
 *                            :1989-2050 and get_process_attributes :1946-1987)
 *   pycd_msd              <- Analysis.compute_msd lag loop
 *                            (PyCD/core.py:2996-3007)
 *   pycd_kmc_batched_*    <- many independent Run.do_kmc_steps trajectories
 *                            (PyCD/core.py:2668 loop over trajectories) in one
 *                            persistent launch
 *
 * Conventions
 *   - Plain pointers and sizes only.  Every buffer is owned by the caller and
 *     may be host or device memory unless the comment says otherwise; the
 *     library detects which with hipPointerGetAttributes.
 *   - Return value: 0 = done, >0 = resumable condition (see PYCD_NEED_*),
 *     <0 = error; the message is then available from pycd_last_error().
 *   - A context owns one HIP stream on one device and must be used from one
 *     host thread at a time.  No C++ exception crosses this ABI.
 *   - All physical quantities are in atomic units (bohr, Hartree, au time)
 *     exactly as the reference holds them after its unit conversions.
 */
#ifndef PYCD_HIP_H
#define PYCD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYCD_OK 0
#define PYCD_NEED_UNIFORMS 1   /* uniform buffer consumed before the end */
#define PYCD_ERR_ARG -1
#define PYCD_ERR_HIP -2
#define PYCD_ERR_NOMEM -3
#define PYCD_ERR_STATE -4

#define PYCD_PREC_F64 0     /* reciprocal sum in FP64 (drop-in)            */
#define PYCD_PREC_F32 1     /* reciprocal sum on FP32 MFMA                 */
#define PYCD_PREC_BF16X3 2  /* three BF16 MFMA products, FP32 accumulation */

typedef struct pycd_ctx pycd_ctx;

/* Context: device selection, stream, scratch.  (No reference counterpart:
 * PyCD has no device state.) */
int pycd_ctx_create(int device, pycd_ctx** out);
void pycd_ctx_destroy(pycd_ctx* ctx);
const char* pycd_last_error(const pycd_ctx* ctx);
/* Compute units of the context's device. */
int pycd_ctx_cu_count(const pycd_ctx* ctx);
/* Symbol present so that loaders can check the ABI revision. */
int pycd_abi_version(void);

/* Device buffers for callers without their own allocator (benchmarks,
 * ctypes users): allocate, free, and copy in any direction (synchronous on
 * the context's stream).  No reference counterpart. */
int pycd_dev_alloc(pycd_ctx* ctx, int64_t bytes, void** out);
int pycd_dev_free(pycd_ctx* ctx, void* p);
int pycd_memcpy(pycd_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* ---------------------------------------------------------------- Ewald -- */
typedef struct {
  int64_t n_kvec;        /* half-space k vectors with |k|^2 < k_cut^2       */
  int64_t n_real_pairs;  /* ordered pairs (incl. diagonal) with r < r_cut   */
  int32_t k_max[3];      /* ceil(k_cut / |b_i|)                             */
  int32_t precision;     /* PYCD_PREC_*                                     */
  double volume;         /* |T0 . (T1 x T2)|                                */
  double ms_sf;          /* structure-factor kernel time                    */
  double ms_real;        /* time of the real-space kernel ahead of the      */
                         /* BF16x3 256 x 256-tile contraction (included in  */
                         /* ms_recip); 0 where the term is in the table     */
                         /* kernel's epilogue                               */
  double ms_recip;       /* reciprocal contraction + epilogue kernel time   */
  double ms_total;       /* device time of the whole table                  */
} pycd_ewald_report;

/* Ewald pair table A (N x N, row-major FP64) with E = q^T A q:
 *   A = A_real/eps + A_recip/eps - sqrt(alpha/pi)/eps I
 * following System.get_precomputed_array (PyCD/core.py:1594-1661).
 *   pos        N x 3 site coordinates (bohr)
 *   cell       3x3 row-major minimum-image cell, rows = system_size[i] *
 *              lattice[i] (PyCD/core.py:296)
 *   trans      3x3 row-major translational matrix used for the volume and
 *              the reciprocal lattice (np.multiply(system_size, lattice),
 *              PyCD/core.py:709-710)
 *   pbc        periodic flags per axis
 *   alpha      the reference's alpha after its unit conversion (used as
 *              eta^2: erfc(sqrt(alpha) r), exp(-k^2/(4 alpha)))
 *   r_cut, k_cut, eps_r   as in the reference, in bohr / bohr^-1
 *   precision  PYCD_PREC_*
 *   table_out  N*N doubles, host or device
 */
int pycd_pair_table(pycd_ctx* ctx, const double* pos, int64_t n,
                    const double cell[9], const double trans[9],
                    const int32_t pbc[3], double alpha, double r_cut,
                    double k_cut, double eps_r, int32_t precision,
                    double* table_out, pycd_ewald_report* report);

/* Tile variant of the table kernels for one precision, kept in the context
 * for its later pycd_pair_table calls.  No counterpart in the reference (it
 * has no kernels): the autotuner (pycd_amd/tuning.py autotune_table) times
 * the variants and sets the fastest.  Every variant gives the same table
 * bit for bit.
 *   PYCD_PREC_F64     1 = 64 x 64 tiles, 2 = 128 x 128 tiles
 *   PYCD_PREC_BF16X3  1 = K staged in slices of 32, 2 = slices of 64,
 *                     3 = slices of 32 in two LDS stages, 4 = 256 x 256
 *                     tiles (8 wavefronts) in two LDS stages
 *   PYCD_PREC_F32     1 (the only one)
 *   0 = the default (FP64: 128 x 128 from 2048 sites on; BF16x3: 4 from
 *   4096 sites on, 2 below). */
int pycd_set_table_variant(pycd_ctx* ctx, int32_t precision, int32_t variant);
/* The variant set for `precision` (0 = default), or -1 for a bad call. */
int pycd_table_variant(const pycd_ctx* ctx, int32_t precision);

/* ------------------------------------------------------------ neighbours -- */
/* N x N x 3 minimum-image vectors pos[j] - pos[i]
 * (Neighbors.get_pairwise_min_image_vector_data, PyCD/core.py:571-594).
 * `out` host or device. */
int pycd_min_image_vectors(pycd_ctx* ctx, const double* pos, int64_t n,
                           const double cell[9], const int32_t pbc[3],
                           double* out);

/* Hop neighbour lists (Neighbors.hop_neighbor_sites, PyCD/core.py:485-569,
 * for all shells of generate_neighbor_list, :596-668, in one pass): for
 * centre c and shell s, the positions k into `targets` with
 * lo[s] < |mic(pos[targets[k]] - pos[centers[c]])| <= hi[s], ascending in k.
 * Outputs (host memory): count[c, s], index[c, s, 0..count), vec[c, s, .., 3]
 * with room for max_per_shell entries; if a shell holds more, the call fails
 * and `count` still holds the true counts.  At most 16 shells. */
int pycd_hop_neighbors(pycd_ctx* ctx, const double* pos, int64_t n,
                       const int64_t* centers, int64_t n_centers,
                       const int64_t* targets, int64_t n_targets,
                       const double cell[9], const int32_t pbc[3],
                       const double* lo, const double* hi, int32_t n_shells,
                       int32_t max_per_shell, int32_t* count, int32_t* index,
                       double* vec);

/* ------------------------------------------------------------- plain KMC -- */
/* Inputs of one trajectory family (one table, one set of hop tables).
 * Per-site hop slots follow the process order of get_process_attributes
 * (PyCD/core.py:1957-1978): hop-distance type, then ascending neighbour.
 * Every carrier's site has exactly `z` slots (uniform neighbour counts). */
typedef struct {
  int64_t n_sites;            /* N                                          */
  int32_t n_carriers;         /* n_c                                        */
  int32_t z;                  /* hop slots per carrier site                 */
  int32_t max_z;              /* row stride of the slot tables              */
  int32_t pad0;
  const double* table;        /* N x N pair table (host or device)          */
  const double* lattice_q;    /* N lattice charges (no carriers)            */
  const int32_t* nbr_site;    /* N x max_z target site per slot             */
  const double* nbr_vec;      /* N x max_z x 3 hop vector per slot (bohr)   */
  const double* nbr_lambda;   /* N x max_z reorganisation energy            */
  const double* nbr_vab;      /* N x max_z electronic coupling              */
  const double* site_energy;  /* N relative site energies                   */
  const double* carrier_q;    /* n_c carrier charges                        */
  double field[3];            /* electric field (au, no conversion)         */
  double nu;                  /* attempt frequency (1 / au time)            */
  double kT;                  /* temperature (Hartree)                      */
  double time_interval;       /* path-step width (au time)                  */
  int64_t n_path;             /* P = int(t_final / time_interval) + 1       */
  /* Optional (all 0: none): the supercell in unit cells and the sites per
   * cell, sites indexed ((x s_y + y) s_z + z) n_unit + e as PyCD indexes
   * them.  With it, the hop loop gathers from the n_unit table rows of
   * cell 0 when the table is above 32 MB (the table of a periodic
   * supercell is translation-invariant; checked entry by entry at
   * creation, to 1e-14 of the table's scale, and not used where the check
   * fails). */
  int32_t cells[3];
  int32_t n_unit;
} pycd_kmc_problem;

typedef struct pycd_kmc pycd_kmc;

/* Uploads the tables, forms phi_lattice = A q_lattice and the self terms
 * A[s,s] - A[s,n] of every slot.  The table pointer may be released by the
 * caller afterwards. */
int pycd_kmc_create(pycd_ctx* ctx, const pycd_kmc_problem* pb,
                    pycd_kmc** out);
void pycd_kmc_destroy(pycd_kmc* s);

/* Replaces the lattice charges and site energies (n_sites doubles each,
 * host or device) for the trajectories begun after this call, and forms
 * phi_lattice and q_lattice^T A q_lattice again; the table stays resident.
 * A doped trajectory has its own of both: dopant sites carry the dopant's
 * charge (PyCD/core.py:2545-2550) and sites within a dopant's shells are
 * shifted in energy (PyCD/core.py:2723-2761).  Ends any running
 * trajectory: call pycd_kmc_begin next. */
int pycd_kmc_set_sites(pycd_kmc* s, const double* lattice_q,
                       const double* site_energy);

/* Table rows the hop loop gathers from: n_unit when the problem gave its
 * cell shape and the table passed the translation check, n_sites
 * otherwise.  No reference counterpart. */
int pycd_kmc_table_rows(const pycd_kmc* s, int64_t* rows);

/* Kernel of the hop loop.  PYCD_KMC_WAVE: one wavefront that keeps the
 * table entries of every carrier at every site the rates need in LDS
 * (n_c <= 64, n_c z <= 256, z <= 8 and the cache within a workgroup's
 * LDS); PYCD_KMC_WORKGROUP: a 256-thread workgroup that gathers them again
 * each hop (any size); PYCD_KMC_AUTO (the default): the wavefront kernel
 * where it applies.  Both give the same results bit for bit.  Asking for
 * PYCD_KMC_WAVE where it does not apply is an argument error.
 * pycd_kmc_variant reports the kernel the next pycd_kmc_advance runs.  No
 * reference counterpart. */
#define PYCD_KMC_AUTO 0
#define PYCD_KMC_WORKGROUP 1
#define PYCD_KMC_WAVE 2
int pycd_kmc_set_variant(pycd_kmc* s, int32_t variant);
int pycd_kmc_variant(const pycd_kmc* s, int32_t* variant);

/* Starts a trajectory: carriers at `occupancy` (n_c sites), clock 0,
 * energy = ewald_neut + q^T A q (PyCD/core.py:2773-2780). */
int pycd_kmc_begin(pycd_kmc* s, const int64_t* occupancy, double ewald_neut);

/* Advances the trajectory, consuming two uniforms per hop (u1 selects the
 * process, u2 advances the clock, PyCD/core.py:2799-2803) from `uniforms`
 * (2*n_pairs doubles, host or device), until the path is full, `max_hops`
 * hops have been made in total (max_hops <= 0: no limit), or the buffer is
 * used up.  Per hop it writes, if the pointers are non-null, the clock after
 * the hop to hop_time[k] and the chosen process index to hop_proc[k]
 * (k = hop index within this call; host or device, n_pairs capacity).
 * Returns PYCD_OK when finished, PYCD_NEED_UNIFORMS when the buffer ran out
 * first.  *hops_out = hops made by this call. */
int pycd_kmc_advance(pycd_kmc* s, const double* uniforms, int64_t n_pairs,
                     int64_t max_hops, double* hop_time, int32_t* hop_proc,
                     int64_t* hops_out);

/* pycd_kmc_advance without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No reference counterpart.
 * The same hop loop and results as pycd_kmc_advance, but every buffer must
 * be device memory, and the log(u2) staging and the status words live in a
 * fixed device array of the library (one per device, 8 MiB, kept until the
 * process ends) instead of session buffers grown on demand: at most 2^20
 * uniform pairs per call (PYCD_ERR_ARG beyond).  Calls on one device share
 * that array: calls that overlap on a device race.  The session itself
 * (pycd_kmc_create, pycd_kmc_begin) is set up beforehand as usual; this
 * call does not update pycd_kmc_last_ms. */
int pycd_kmc_advance_static(pycd_kmc* s, const double* uniforms,
                            int64_t n_pairs, int64_t max_hops,
                            double* hop_time, int32_t* hop_proc,
                            int64_t* hops_out);

/* Device time (ms, HIP events on the context's stream) of the kernel of
 * the last pycd_kmc_advance call. */
double pycd_kmc_last_ms(const pycd_kmc* s);

/* Copies out the trajectory outputs (any pointer may be null):
 *   unwrapped  P x 3 n_c   (unwrapped_traj.npy)
 *   energy     P           (energy_traj.npy)
 *   delg_0     P           (delG0_traj.npy)
 *   drift      n_c x 3     sum of hop vector * rate (PyCD/core.py:2822-2825)
 *   occupancy  n_c         current sites
 *   scalars    [t, energy, start_path_index, hops]                          */
int pycd_kmc_read(pycd_kmc* s, double* unwrapped, double* energy,
                  double* delg_0, double* drift, int64_t* occupancy,
                  double scalars[4]);

/* ------------------------------------------------------------------ MSD -- */
/* Squared displacement per lag (PyCD/core.py:2996-3007):
 *   sd[t, tau, c] = mean_{t0 < P - tau} |pos[t, t0 + tau, c] - pos[t, t0, c]|^2
 * for tau in [1, n_lags), sd[t, 0, c] = 0.
 *   pos  n_traj x P x n_c x 3 (already in output units), host or device
 *   sd   n_traj x n_lags x n_c                                              */
int pycd_msd(pycd_ctx* ctx, const double* pos, int64_t n_traj,
             int64_t n_path, int64_t n_carriers, int64_t n_lags, double* sd);

/* --------------------------------------------------------- batched KMC -- */
/* Many independent trajectories (the loop over trajectories of
 * Run.do_kmc_steps, PyCD/core.py:2668, each running the loop body of
 * PyCD/core.py:2787-2861) in one persistent launch.  A family is one
 * supercell: its table and hop slots are shared by all its trajectories.
 * Every trajectory draws its uniforms from its own CPython MT19937 state on
 * the device (4 words per hop, as random.random() twice).
 *
 * A family must be a periodic supercell of identical unit cells with the
 * sites numbered cell-major, site = ((cx cells[1] + cy) cells[2] + cz)
 * n_unit + sub (Material.generate_sites, PyCD/core.py:154-241): the library
 * checks that every per-site value, hop slot and (with table_rows = N)
 * table entry is the cell-0 one moved by the site's cell, and keeps only
 * the table rows of cell 0's sites on the device (n_unit x N).  N must be
 * below 2^24 and cells[1] cells[2] n_unit below 65536. */
typedef struct {
  int64_t n_sites;            /* N = cells[0] cells[1] cells[2] n_unit      */
  int32_t z;                  /* hop slots per hopping site (1..7)          */
  int32_t max_z;              /* row stride of the slot tables              */
  int32_t cells[3];           /* supercell size in unit cells (1..255 each) */
  int32_t n_unit;             /* sites per unit cell (1..255, n_unit z<=256)*/
  int64_t table_rows;         /* N: full table, checked entry by entry;
                                 n_unit: the rows of cell 0's sites only    */
  const double* table;        /* table_rows x N pair table                  */
  const double* lattice_q;    /* N lattice charges                          */
  const int32_t* nbr_site;    /* N x max_z                                  */
  const double* nbr_vec;      /* N x max_z x 3                              */
  const double* nbr_lambda;   /* N x max_z                                  */
  const double* nbr_vab;      /* N x max_z                                  */
  const double* site_energy;  /* N                                          */
} pycd_kmc_family;

typedef struct {
  int32_t family;             /* index into the families                    */
  int32_t n_carriers;         /* 1..64                                      */
  double carrier_q;           /* charge of every carrier                    */
  double field[3];            /* au                                         */
  double kT;                  /* Hartree                                    */
  double time_interval;       /* path-step width (au time)                  */
  double ewald_neut;          /* background term of the energy              */
  int64_t n_path;             /* path rows                                  */
  int64_t max_hops;           /* stop after this many hops (<= 0: time)     */
} pycd_kmc_traj;

typedef struct {
  double t;                   /* clock after the last hop                   */
  double energy;              /* energy after the last hop                  */
  int64_t hops;
  int64_t start;              /* path rows written: [0, start)              */
  int32_t no_select;          /* hops where rounding left no cum > u1       */
  int32_t done;
  double busy_s;              /* device time its wavefront spent on it      */
  int32_t simd;               /* SIMD that ran it last: XCC << 10 | SE << 7 |
                                 SH << 6 | CU << 2 | SIMD (from HW_ID)       */
  int32_t pad0;
} pycd_kmc_traj_result;

typedef struct pycd_batch pycd_batch;

/* Uploads the families; table_precision / rate_precision are PYCD_PREC_F64
 * or PYCD_PREC_F32 (the prefix sum, selection and clock are FP64 always). */
int pycd_batch_create(pycd_ctx* ctx, const pycd_kmc_family* fams,
                      int32_t n_fam, double nu, int32_t table_precision,
                      int32_t rate_precision, pycd_batch** out);
void pycd_batch_destroy(pycd_batch* b);

/* Trajectories, their concatenated initial occupancies, their MT states
 * (n_traj x 625 words: the 624-word key and the position, as in
 * random.getstate()), an optional launch order (a permutation; NULL: as
 * given), optional per-trajectory MSD lag counts (NULL: no MSD) with the
 * length scale applied to positions first, and waves per CU (<= 0: 8). */
int pycd_batch_setup(pycd_batch* b, const pycd_kmc_traj* trajs,
                     int64_t n_traj, const int64_t* occupancy,
                     const uint32_t* mt_states, const int32_t* order,
                     const int64_t* msd_lags, double msd_scale,
                     int32_t waves_per_cu);

/* Hand-over of running trajectories to idle wavefronts in a launch with
 * SIMD bins (default on).  SIMDs of one chip run the same work at speeds up
 * to ~8% apart; with it, a trajectory with hops to spare moves, at a hop
 * boundary and with its exact state, to a wavefront whose SIMD ran out of
 * work.  Results are the same either way.  No reference counterpart (the
 * reference runs trajectories one after another, PyCD/core.py:2668). */
int pycd_batch_set_handover(pycd_batch* b, int32_t on);

/* Whether launches read the families' hop differences (default on; kept
 * for each family whose n_unit z x N FP64 rows stay below 4 GiB): the
 * change of every slot potential when a carrier hops is then one table
 * gather instead of two, the same number either way.  Off, or for a family
 * without them, the two entries are gathered.  No reference counterpart
 * (the reference sums the potential from the full table each hop,
 * PyCD/core.py:1989-2050). */
int pycd_batch_set_hop_differences(pycd_batch* b, int32_t on);

/* Runs every trajectory (and the MSD reduction), optionally through a
 * captured hipGraph; *elapsed_ms = device time between events around it. */
int pycd_batch_run(pycd_batch* b, int32_t use_graph, double* elapsed_ms);

/* Adds a family to an existing batch (index returned), or frees one that
 * no busy slot uses.  Used by the serving loop's table cache. */
int pycd_batch_add_family(pycd_batch* b, const pycd_kmc_family* f,
                          int32_t* index);
int pycd_batch_drop_family(pycd_batch* b, int32_t index);

/* Serving (continuous batching).  A pool of n_slots trajectory slots with a
 * path-row arena of rows_capacity doubles; every slot starts idle.
 * pycd_batch_admit fills an idle slot (its path rows at row_off in the
 * arena, which must not overlap a busy slot's); pycd_batch_step advances
 * every admitted, unfinished slot by at most slice_hops hops in one launch
 * (state is saved on the device between steps, so a trajectory's result
 * does not depend on how it is sliced or what else runs); a finished or
 * cancelled slot is handed back with pycd_batch_release.  sub_max (<= 256)
 * bounds n_unit z of every family the pool will run (its per-wavefront LDS
 * holds one entry per sublattice slot). */
int pycd_batch_pool(pycd_batch* b, int64_t n_slots, int32_t nc_max,
                    int32_t z_max, int32_t sub_max, int64_t rows_capacity,
                    int32_t waves_per_cu);
int pycd_batch_admit(pycd_batch* b, int64_t slot, const pycd_kmc_traj* traj,
                     const int64_t* occupancy, const uint32_t* mt_state,
                     int64_t row_off);
int pycd_batch_release(pycd_batch* b, int64_t slot);
int pycd_batch_step(pycd_batch* b, int64_t slice_hops, double* elapsed_ms);
/* Carrier-averaged MSD of the given pool slots (lags[k] lags each, positions
 * scaled by `scale` first), concatenated. */
int pycd_batch_slot_msd(pycd_batch* b, const int64_t* slots, int64_t n,
                        const int64_t* lags, double scale, double* out);

/* Device time of the KMC kernel alone (no MSD, no resets) in the last
 * pycd_batch_run without a graph, or the last pycd_batch_step. */
int pycd_batch_kernel_ms(const pycd_batch* b, double* ms);
/* Hand-overs claimed in the last launch (0 without SIMD bins or with
 * hand-over off).  No reference counterpart. */
int pycd_batch_handovers(pycd_batch* b, int64_t* n);
/* Wavefronts of the KMC kernel that one CU holds at once for the current
 * set-up (registers and LDS); a grid beyond that many per CU queues. */
int pycd_batch_resident(const pycd_batch* b, int32_t* per_cu);

/* Per-trajectory results and drift sums (n_traj x 64 x 3, carrier-major). */
int pycd_batch_results(pycd_batch* b, pycd_kmc_traj_result* results,
                       double* drift);
/* Path rows of one trajectory (n_path x 3 n_c; rows never written are 0). */
int pycd_batch_rows(pycd_batch* b, int64_t traj, double* rows);
/* Carrier-averaged MSD per trajectory, concatenated (sum of lag counts). */
int pycd_batch_msd(pycd_batch* b, double* msd);

#ifdef __cplusplus
}
#endif

#endif /* PYCD_HIP_H */

"""Serving form: continuous batching of KMC jobs on one GPU.

Jobs (one trajectory each: a lattice size, a carrier count, a hop budget and
a seed) arrive over time.  `KmcServer` keeps a pool of trajectory slots in
one `pycd_batch` and loops:

  1. admit waiting jobs, first come first served, while a slot is free, the
     job's path rows fit the device arena, and its pair table is cached or
     fits the table budget (tables are built on the GPU on first use and
     kept resident for repeat sizes; least-recently-used tables with no
     running job are evicted when the budget needs room);
  2. advance every admitted job by at most `slice_hops` hops in one launch
     of the batched kernel (jobs of different sizes share the launch);
  3. retire the jobs that finished: copy their results, reduce their MSD on
     the device, free their slot and rows.

The kernel saves each trajectory's complete state between launches, so a
job's result is the same whatever else shares its launches and however it
is sliced (tests/test_gpu_serve.py checks this bit for bit).

Failure handling, all from the host between launches (every kernel exits
normally): `cancel(job)` stops a pending or running job and frees its slot
and rows; `cancel(job, requeue=True)` starts it over from its seed; a
launch whose device time exceeds `watchdog_ms` is treated as a timed-out
slice, and every job in it is re-queued from the start, up to
`max_timeouts` times per job (then the job is rejected, so a slice that
always overruns cannot keep the loop busy forever).  A job's stream and
time grid are fixed by its parameters, so a re-run gives the same result.
"""
import ctypes as C
import time
from collections import OrderedDict, deque

import numpy as np

from . import _native, constants
from ._native import lib, _check, _ptr, TrajC, TrajResultC, FamilyC
from .batch import Family, expected_rate_per_carrier
from .rng import seeded_starts


class RowArena:
    """First-fit allocator over [0, capacity) doubles of the path arena."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.free = [(0, self.capacity)]

    def alloc(self, n):
        n = int(n)
        for k, (lo, size) in enumerate(self.free):
            if size >= n:
                if size == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (lo + n, size - n)
                return lo
        return None

    def release(self, lo, n):
        self.free.append((int(lo), int(n)))
        self.free.sort()
        merged = []
        for a, s in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + s)
            else:
                merged.append((a, s))
        self.free = merged

    @property
    def used(self):
        return self.capacity - sum(s for _, s in self.free)


class Job:
    def __init__(self, job_id, edge, n_carriers, hops, seed, temp=300.0,
                 field_vpm=0.0, field_dir=(0, 0, 1), arrival=0.0):
        self.id = job_id
        self.edge = int(edge)
        self.n_carriers = int(n_carriers)
        self.hops = int(hops)
        self.seed = seed
        self.temp = float(temp)
        self.field_vpm = float(field_vpm)
        self.field_dir = np.asarray(field_dir, dtype=np.float64)
        self.arrival = float(arrival)
        self.admitted = None
        self.finished = None
        self.result = None
        self.slot = None
        self.rows_at = None
        self.n_path = None
        self.timeouts = 0


class KmcServer:
    def __init__(self, family_builder, n_slots=2048, nc_max=32, z_max=4,
                 rows_capacity=1 << 28, table_budget_bytes=32 << 30,
                 slice_hops=2000, rate_precision="f32", path_rows=200,
                 msd_lags=100, waves_per_cu=8, nu=None, watchdog_ms=None,
                 sub_max=64, max_timeouts=2):
        """family_builder(edge) -> (Run, table) for that lattice size.
        sub_max bounds sites-per-cell x hop slots of the families served;
        a job in more than `max_timeouts` timed-out slices is rejected."""
        self.build_family = family_builder
        self.slice_hops = int(slice_hops)
        self.path_rows = int(path_rows)
        # a job's time grid spans ~4x the estimated duration of its hop
        # budget (the zero-interaction rate estimate runs fast for many
        # carriers, and a full path ends the job)
        self.n_path = 4 * self.path_rows + 1
        self.msd_lags = int(msd_lags)
        self.table_budget = int(table_budget_bytes)
        self.nc_max = int(nc_max)
        self.ctx = None             # the device context, with the batch
        self.families = {}          # edge -> (index, Family)
        self.lru = OrderedDict()    # edge -> last use
        self.family_bytes = {}
        self._nu = nu
        self.h = None
        self.n_slots = int(n_slots)
        self.z_max = int(z_max)
        self.sub_max = int(sub_max)
        self.rows = RowArena(rows_capacity)
        self.waves_per_cu = int(waves_per_cu)
        self.rate_precision = rate_precision
        self.free_slots = deque(range(self.n_slots))
        self.active = {}            # slot -> job
        self.pending = deque()
        self.done = []
        self.rejected = []
        self.launch_ms = []
        self.evictions = 0
        self.watchdog_ms = watchdog_ms
        self.max_timeouts = int(max_timeouts)
        self.timeouts = 0
        self.requeued = 0
        # device footprint per edge, and the last family built only to
        # measure it (taken by the admission that follows, not built again)
        self._footprint = {}
        self._staged = None

    # -- device objects ----------------------------------------------------
    def _ensure_batch(self, nu):
        if self.h is not None:
            return
        self.ctx = _native.context()
        h = C.c_void_p()
        _check(self.ctx, lib().pycd_batch_create(
            self.ctx, None, 0, float(nu), _native.PREC["f64"],
            _native.PREC[self.rate_precision], C.byref(h)),
            "pycd_batch_create")
        self.h = h
        _check(self.ctx, lib().pycd_batch_pool(
            self.h, self.n_slots, self.nc_max, self.z_max, self.sub_max,
            self.rows.capacity, self.waves_per_cu), "pycd_batch_pool")

    def close(self):
        if self.h is not None:
            lib().pycd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- table cache -----------------------------------------------------------
    def _busy_edges(self):
        return {j.edge for j in self.active.values()}

    def _family_for(self, edge):
        """Family index for `edge`, building (and evicting) as needed.
        Returns None when the table does not fit beside the busy ones."""
        if edge in self.families:
            self.lru.move_to_end(edge)
            return self.families[edge][0]
        if self._staged is not None and self._staged[0] == edge:
            fam = self._staged[1]
        else:
            fam = Family(*self.build_family(edge))
        self._staged = (edge, fam)
        need = fam.device_bytes("f64")
        busy = self._busy_edges()
        while sum(self.family_bytes.values()) + need > self.table_budget:
            victims = [e for e in self.lru if e not in busy]
            if not victims:
                return None
            self._drop(victims[0])
        self._ensure_batch(fam.nu if self._nu is None else self._nu)
        idx = C.c_int32()
        _check(self.ctx, lib().pycd_batch_add_family(self.h, fam.cstruct(),
                                                     C.byref(idx)),
               "pycd_batch_add_family")
        fam.table = None  # the device holds it now
        self._staged = None
        self.families[edge] = (idx.value, fam)
        self.family_bytes[edge] = need
        self.lru[edge] = time.perf_counter()
        return idx.value

    def _drop(self, edge):
        idx, _ = self.families.pop(edge)
        _check(self.ctx, lib().pycd_batch_drop_family(self.h, idx),
               "pycd_batch_drop_family")
        del self.lru[edge]
        del self.family_bytes[edge]
        self.evictions += 1

    # -- jobs --------------------------------------------------------------------
    def submit(self, job):
        self.pending.append(job)

    def _reject(self, job, why):
        job.result = {"error": why, "hops": 0}
        job.finished = time.perf_counter()
        self.rejected.append(job)

    def _never_fits(self, job):
        """Reason the job can never be admitted, or None."""
        if job.n_carriers > self.nc_max:
            return f"{job.n_carriers} carriers > pool limit {self.nc_max}"
        if self.n_path * 3 * job.n_carriers > self.rows.capacity:
            return "path rows exceed the arena"
        if job.edge in self.families:
            return None
        need = self.table_bytes(job.edge)
        if need > self.table_budget:
            return f"pair table ({need} B) exceeds the table budget"
        return None

    def table_bytes(self, edge):
        """Device bytes of the family's compact form (cell-0 table rows),
        measured once per edge."""
        if edge not in self._footprint:
            if self._staged is not None and self._staged[0] == edge:
                fam = self._staged[1]
            else:
                fam = Family(*self.build_family(edge))
                self._staged = (edge, fam)
            self._footprint[edge] = fam.device_bytes("f64")
        return self._footprint[edge]

    def _try_admit(self, job):
        if not self.free_slots:
            return False
        n_path = self.n_path
        need = n_path * 3 * job.n_carriers
        at = self.rows.alloc(need)
        if at is None:
            return False
        fidx = self._family_for(job.edge)
        if fidx is None:
            self.rows.release(at, need)
            return False
        fam = self.families[job.edge][1]
        occs, words = seeded_starts(job.seed, 1, [fam.hop_sites],
                                    [job.n_carriers])
        kT = job.temp * constants.K2AUTEMP
        d = job.field_dir
        field = (job.field_vpm * constants.VPM2AUFIELD * d / np.linalg.norm(d)
                 if job.field_vpm else np.zeros(3))
        rate = job.n_carriers * expected_rate_per_carrier(fam, kT, field)
        t = TrajC()
        t.family = fidx
        t.n_carriers = job.n_carriers
        t.carrier_q = -1.0
        for a in range(3):
            t.field[a] = float(field[a])
        t.kT = kT
        t.time_interval = job.hops / rate / self.path_rows
        t.ewald_neut = 0.0
        t.n_path = n_path
        t.max_hops = job.hops
        slot = self.free_slots.popleft()
        occ = np.asarray(occs[0], dtype=np.int64)
        w = np.ascontiguousarray(words[0], dtype=np.uint32)
        _check(self.ctx, lib().pycd_batch_admit(self.h, slot, C.byref(t),
                                                _ptr(occ), _ptr(w), at),
               "pycd_batch_admit")
        job.slot, job.rows_at, job.n_path = slot, at, n_path
        job.time_interval = t.time_interval
        job.admitted = time.perf_counter()
        self.active[slot] = job
        return True

    def admit_pending(self):
        """Admit waiting jobs in arrival order while they fit.  A job that
        could never fit is rejected (result["error"]); a job that fits an
        empty server but not the current one waits for running jobs."""
        admitted = 0
        while self.pending:
            job = self.pending[0]
            why = self._never_fits(job)
            if why is not None:
                self.pending.popleft()
                self._reject(job, why)
                continue
            if not self._try_admit(job):
                if not self.active:
                    # nothing running will free room: cannot happen for a job
                    # that passed _never_fits, but never spin on it
                    self.pending.popleft()
                    self._reject(job, "does not fit an idle server")
                    continue
                break
            self.pending.popleft()
            admitted += 1
        return admitted

    def cancel(self, job, requeue=False):
        """Stop a pending or running job and free its slot and rows.  With
        `requeue` it goes back to the head of the queue and starts over;
        otherwise its result is {"error": "cancelled"}."""
        if job in self.pending:
            self.pending.remove(job)
        elif job.slot is not None and self.active.get(job.slot) is job:
            s = job.slot
            del self.active[s]
            self._release_slot(s)
            self.rows.release(job.rows_at, job.n_path * 3 * job.n_carriers)
            self.free_slots.append(s)
        else:
            return False
        job.slot = job.rows_at = None
        if requeue:
            self.requeued += 1
            self.pending.appendleft(job)
        else:
            job.result = {"error": "cancelled", "hops": 0}
            job.finished = time.perf_counter()
            self.rejected.append(job)
        return True

    def _release_slot(self, slot):
        _check(self.ctx, lib().pycd_batch_release(self.h, slot),
               "pycd_batch_release")

    def _launch_slice(self):
        """One pycd_batch_step launch; returns its device time in ms."""
        ms = C.c_double(0.0)
        _check(self.ctx, lib().pycd_batch_step(self.h, self.slice_hops,
                                               C.byref(ms)),
               "pycd_batch_step")
        return ms.value

    def step(self):
        """One slice over every admitted job; retire the finished ones."""
        if not self.active:
            return []
        ms = self._launch_slice()
        self.launch_ms.append(ms)
        if self.watchdog_ms is not None and ms > self.watchdog_ms:
            # a timed-out slice: its jobs start over, newest first so that
            # the queue keeps their arrival order; a job past its limit of
            # timed-out slices is rejected instead
            self.timeouts += 1
            for job in sorted(self.active.values(),
                              key=lambda j: -j.arrival):
                job.timeouts += 1
                if job.timeouts > self.max_timeouts:
                    self.cancel(job)
                    job.result = {"error": f"timed out in {job.timeouts} "
                                           f"slices", "hops": 0}
                else:
                    self.cancel(job, requeue=True)
            return []
        res = (TrajResultC * self.n_slots)()
        _check(self.ctx, lib().pycd_batch_results(self.h, res, None),
               "pycd_batch_results")
        finished = [s for s in self.active if res[s].done]
        if finished:
            self._retire(finished, res)
        return finished

    def _retire(self, slots, res):
        now = time.perf_counter()
        arr = np.asarray(slots, dtype=np.int64)
        lags = np.full(len(slots), self.msd_lags, dtype=np.int64)
        msd = np.empty(int(lags.sum()))
        _check(self.ctx, lib().pycd_batch_slot_msd(
            self.h, _ptr(arr), len(slots), _ptr(lags),
            1.0 / constants.ANG2BOHR, _ptr(msd)), "pycd_batch_slot_msd")
        for k, s in enumerate(slots):
            job = self.active.pop(s)
            r = res[s]
            rows = np.empty((job.n_path, 3 * job.n_carriers))
            _check(self.ctx, lib().pycd_batch_rows(self.h, s, _ptr(rows)),
                   "pycd_batch_rows")
            job.result = {"t": r.t, "energy": r.energy, "hops": r.hops,
                          "start": r.start, "no_select": r.no_select,
                          "rows": rows,
                          "msd": msd[k * self.msd_lags:(k + 1) * self.msd_lags]}
            job.finished = now
            self._release_slot(s)
            self.rows.release(job.rows_at, job.n_path * 3 * job.n_carriers)
            self.free_slots.append(s)
            self.done.append(job)

    def drain(self):
        while self.active or self.pending:
            self.admit_pending()
            self.step()

    def serve(self, jobs, log=None):
        """Real-time loop over jobs with `arrival` offsets (s) from now."""
        jobs = sorted(jobs, key=lambda j: j.arrival)
        t0 = time.perf_counter()
        for j in jobs:
            j.arrival_abs = t0 + j.arrival
        k = 0
        while k < len(jobs) or self.active or self.pending:
            now = time.perf_counter()
            while k < len(jobs) and jobs[k].arrival_abs <= now:
                self.submit(jobs[k])
                k += 1
            self.admit_pending()
            if not self.active:
                if k < len(jobs):
                    time.sleep(max(0.0, jobs[k].arrival_abs
                                   - time.perf_counter()))
                continue
            self.step()
        wall = time.perf_counter() - t0
        ok = [j for j in jobs if "error" not in j.result]
        lat = np.array([j.finished - j.arrival_abs for j in ok])
        hops = float(sum(j.result["hops"] for j in ok))
        return {"jobs": len(ok), "rejected": len(jobs) - len(ok),
                "wall_s": wall, "hops": hops,
                "hops_per_s": hops / wall,
                "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_latency_ms": float(lat.mean() * 1e3),
                "launches": len(self.launch_ms),
                "device_ms": float(np.sum(self.launch_ms)),
                "table_evictions": self.evictions,
                "timeouts": self.timeouts, "requeued": self.requeued}


def poisson_jobs(n_jobs, rate=200.0, edges=(4, 12), nc=(1, 32),
                 hops=(1e5, 1e6), seed=5):
    """The serving workload: Poisson arrivals, uniform edges and carriers,
    log-uniform hop budgets."""
    g = np.random.default_rng(seed)
    t = np.cumsum(g.exponential(1.0 / rate, size=n_jobs))
    out = []
    for k in range(n_jobs):
        out.append(Job(k, g.integers(edges[0], edges[1] + 1),
                       g.integers(nc[0], nc[1] + 1),
                       int(np.exp(g.uniform(np.log(hops[0]),
                                            np.log(hops[1])))),
                       seed=1000 + k, arrival=float(t[k])))
    return out

"""Batched and served forms of the KMC path (host side of pycd_batch_*).

`KmcBatch` holds one or more lattice families (one pair table and hop-slot
table each, shared by every trajectory of that size) on the device and runs
any number of independent trajectories of those families in one persistent
kernel launch, optionally replayed through a captured hipGraph.

A family is a periodic supercell: the library checks that the table and
the hop slots are invariant under cell translations and keeps only the
table rows of the sites of cell 0 (n_unit x N) on the device.
"""
import ctypes as C

import numpy as np

from . import _native
from ._native import (lib, _check, _ptr, _f64, PREC, FamilyC, TrajC,
                      TrajResultC)

MAX_CARRIERS = 64


class Family:
    """Host arrays of one lattice size (from a Run and its table).

    `table` is the full N x N pair table (checked entry by entry against
    its cell-0 rows when the family is uploaded) or only the n_unit rows of
    the sites of cell 0.
    """

    def __init__(self, run, table):
        self.run = run
        self.cells = [int(v) for v in run.system_size]
        n_cells = int(np.prod(self.cells))
        table = np.asarray(table)
        n_sites = table.shape[1]
        if n_sites % n_cells:
            raise ValueError("site count is not a multiple of the cell count")
        self.n_unit = n_sites // n_cells
        if table.shape[0] not in (n_sites, self.n_unit):
            raise ValueError("table must be N x N or n_unit x N")
        self.n_sites = n_sites
        self.table = _f64(table)
        self.lattice_q = _f64(run.base_charge_config()[:, 0])
        self.nbr_site = np.ascontiguousarray(run.nbr_site, dtype=np.int32)
        self.nbr_vec = _f64(run.nbr_vec)
        self.nbr_lambda = _f64(run.nbr_lambda)
        self.nbr_vab = _f64(run.nbr_vab)
        self.site_energy = _f64(run.system_relative_energies)
        counts = set(int(c) for c in run.nbr_count if c > 0)
        if len(counts) != 1:
            raise ValueError("non-uniform hop slot counts")
        self.z = counts.pop()
        self.hop_sites = list(run.hop_site_lists[run.species_type_list[0]])
        self.nu = run.material.vn
        self.volume = run.system.system_volume
        self.alpha = run.system.alpha

    def device_bytes(self, table_precision="f64"):
        """Device footprint of the family's compact form: the cell-0 rows,
        the hop differences (n_unit z rows of FP64, kept while they stay
        below 4 GiB), per-sublattice records and per-site slot words."""
        rows = self.n_unit * self.n_sites
        per = 12 if table_precision == "f32" else 8
        diff = 8 * self.n_unit * self.z * self.n_sites
        if diff >= 1 << 32:
            diff = 0
        return (rows * per + diff + self.n_unit * (8 + 72 * self.z)
                + 4 * self.n_sites)

    def compact(self):
        """The same family with only the table rows of cell 0's sites."""
        out = object.__new__(Family)
        out.__dict__.update(self.__dict__)
        out.table = np.ascontiguousarray(self.table[:self.n_unit])
        return out

    def cstruct(self):
        f = FamilyC()
        f.n_sites = self.n_sites
        f.z = self.z
        f.max_z = self.nbr_site.shape[1]
        for a in range(3):
            f.cells[a] = self.cells[a]
        f.n_unit = self.n_unit
        f.table_rows = self.table.shape[0]
        for name in ("table", "lattice_q", "nbr_site", "nbr_vec",
                     "nbr_lambda", "nbr_vab", "site_energy"):
            setattr(f, name, getattr(self, name).ctypes.data)
        return f


class KmcBatch:
    def __init__(self, families, table_precision="f64", rate_precision="f64"):
        self.families = list(families)
        nus = {f.nu for f in self.families}
        if len(nus) != 1:
            raise ValueError("all families must share the attempt frequency")
        self.nu = nus.pop()
        self.ctx = _native.context()
        arr = (FamilyC * len(self.families))(
            *[f.cstruct() for f in self.families])
        h = C.c_void_p()
        rc = lib().pycd_batch_create(self.ctx, arr, len(self.families),
                                     float(self.nu), PREC[table_precision],
                                     PREC[rate_precision], C.byref(h))
        _check(self.ctx, rc, "pycd_batch_create")
        self.h = h
        self.n_traj = 0

    def close(self):
        if getattr(self, "h", None):
            lib().pycd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self, trajs, occupancies, mt_words, order=None, msd_lags=None,
              msd_scale=1.0, waves_per_cu=8):
        """trajs: list of dicts with family, n_carriers, carrier_q, field,
        kT, time_interval, ewald_neut, n_path, max_hops."""
        n = len(trajs)
        arr = (TrajC * n)()
        for k, t in enumerate(trajs):
            c = arr[k]
            c.family = int(t["family"])
            c.n_carriers = int(t["n_carriers"])
            c.carrier_q = float(t.get("carrier_q", -1.0))
            fld = t.get("field", (0.0, 0.0, 0.0))
            for a in range(3):
                c.field[a] = float(fld[a])
            c.kT = float(t["kT"])
            c.time_interval = float(t["time_interval"])
            c.ewald_neut = float(t.get("ewald_neut", 0.0))
            c.n_path = int(t["n_path"])
            c.max_hops = int(t.get("max_hops", 0) or 0)
        occ = np.ascontiguousarray(np.concatenate(
            [np.asarray(o, dtype=np.int64) for o in occupancies]))
        words = np.ascontiguousarray(mt_words, dtype=np.uint32)
        assert words.shape == (n, 625)
        ordp = None
        if order is not None:
            order = np.ascontiguousarray(order, dtype=np.int32)
            ordp = _ptr(order)
        lags = None
        if msd_lags is not None:
            lags = np.ascontiguousarray(msd_lags, dtype=np.int64)
        rc = lib().pycd_batch_setup(self.h, arr, n, _ptr(occ), _ptr(words),
                                    ordp, _ptr(lags) if lags is not None
                                    else None, float(msd_scale),
                                    int(waves_per_cu))
        _check(self.ctx, rc, "pycd_batch_setup")
        self.trajs = trajs
        self.n_traj = n
        self.msd_lags = lags
        return self

    def set_handover(self, on):
        """Hand running trajectories over to idle wavefronts (default on
        for launches with SIMD bins); results are the same either way."""
        _check(self.ctx, lib().pycd_batch_set_handover(self.h, 1 if on else 0),
               "pycd_batch_set_handover")

    def set_hop_differences(self, on):
        """Read the families' hop differences (default on): one gather per
        slot row and hop instead of two; results are the same either way."""
        _check(self.ctx, lib().pycd_batch_set_hop_differences(
            self.h, 1 if on else 0), "pycd_batch_set_hop_differences")

    def run(self, use_graph=False):
        ms = C.c_double(0.0)
        _check(self.ctx, lib().pycd_batch_run(self.h, 1 if use_graph else 0,
                                              C.byref(ms)), "pycd_batch_run")
        return ms.value

    def kernel_ms(self):
        """Device time of the KMC kernel alone in the last eager run."""
        ms = C.c_double(0.0)
        _check(self.ctx, lib().pycd_batch_kernel_ms(self.h, C.byref(ms)),
               "pycd_batch_kernel_ms")
        return ms.value

    def handovers(self):
        """Trajectories handed over to idle wavefronts in the last launch."""
        n = C.c_int64(0)
        _check(self.ctx, lib().pycd_batch_handovers(self.h, C.byref(n)),
               "pycd_batch_handovers")
        return n.value

    def resident_per_cu(self):
        """Wavefronts of the KMC kernel one CU holds for this set-up."""
        n = C.c_int32(0)
        _check(self.ctx, lib().pycd_batch_resident(self.h, C.byref(n)),
               "pycd_batch_resident")
        return n.value

    def results(self):
        res = (TrajResultC * self.n_traj)()
        drift = np.empty((self.n_traj, MAX_CARRIERS, 3))
        _check(self.ctx, lib().pycd_batch_results(self.h, res, _ptr(drift)),
               "pycd_batch_results")
        out = {k: np.array([getattr(r, k) for r in res])
               for k in ("t", "energy", "hops", "start", "no_select", "done",
                         "busy_s", "simd")}
        out["drift"] = drift
        return out

    def rows(self, t):
        tr = self.trajs[t]
        rows = np.empty((int(tr["n_path"]), 3 * int(tr["n_carriers"])))
        _check(self.ctx, lib().pycd_batch_rows(self.h, int(t), _ptr(rows)),
               "pycd_batch_rows")
        return rows

    def msd(self):
        total = int(self.msd_lags.sum())
        out = np.empty(max(total, 1))
        _check(self.ctx, lib().pycd_batch_msd(self.h, _ptr(out)),
               "pycd_batch_msd")
        parts, k = [], 0
        for m in self.msd_lags:
            parts.append(out[k:k + int(m)])
            k += int(m)
        return parts


def expected_rate_per_carrier(family, kT, field=(0.0, 0.0, 0.0)):
    """Zero-dG0 total hop rate of one carrier at a representative site."""
    run = family.run
    s = family.hop_sites[0]
    lam = run.nbr_lambda[s, :family.z]
    vab = run.nbr_vab[s, :family.z]
    shift = 0.5 * (run.nbr_vec[s, :family.z] @ np.asarray(field))
    gs = lam / 4 - vab - shift
    return float(np.sum(family.nu * np.exp(-gs / kT)))

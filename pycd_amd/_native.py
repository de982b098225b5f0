"""ctypes binding of libpycd_hip.so (declared in include/pycd_hip.h).

The library is required: if it is missing or fails to load, every entry
point raises.  There is no CPU fallback on this path.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get(
    "PYCD_LIB", Path(__file__).resolve().parent / "lib" / "libpycd_hip.so"))

PREC = {"f64": 0, "f32": 1, "bf16x3": 2}
KMC_VARIANT = {"auto": 0, "workgroup": 1, "wave": 2}
OK, NEED_UNIFORMS = 0, 1

_lib = None
_ctx = {}


class NativeError(RuntimeError):
    pass


class EwaldReport(C.Structure):
    _fields_ = [("n_kvec", C.c_int64), ("n_real_pairs", C.c_int64),
                ("k_max", C.c_int32 * 3), ("precision", C.c_int32),
                ("volume", C.c_double), ("ms_sf", C.c_double),
                ("ms_real", C.c_double), ("ms_recip", C.c_double),
                ("ms_total", C.c_double)]


class KmcProblemC(C.Structure):
    _fields_ = [("n_sites", C.c_int64), ("n_carriers", C.c_int32),
                ("z", C.c_int32), ("max_z", C.c_int32), ("pad0", C.c_int32),
                ("table", C.c_void_p), ("lattice_q", C.c_void_p),
                ("nbr_site", C.c_void_p), ("nbr_vec", C.c_void_p),
                ("nbr_lambda", C.c_void_p), ("nbr_vab", C.c_void_p),
                ("site_energy", C.c_void_p), ("carrier_q", C.c_void_p),
                ("field", C.c_double * 3), ("nu", C.c_double),
                ("kT", C.c_double), ("time_interval", C.c_double),
                ("n_path", C.c_int64), ("cells", C.c_int32 * 3),
                ("n_unit", C.c_int32)]


class FamilyC(C.Structure):
    _fields_ = [("n_sites", C.c_int64), ("z", C.c_int32),
                ("max_z", C.c_int32), ("cells", C.c_int32 * 3),
                ("n_unit", C.c_int32), ("table_rows", C.c_int64),
                ("table", C.c_void_p),
                ("lattice_q", C.c_void_p), ("nbr_site", C.c_void_p),
                ("nbr_vec", C.c_void_p), ("nbr_lambda", C.c_void_p),
                ("nbr_vab", C.c_void_p), ("site_energy", C.c_void_p)]


class TrajC(C.Structure):
    _fields_ = [("family", C.c_int32), ("n_carriers", C.c_int32),
                ("carrier_q", C.c_double), ("field", C.c_double * 3),
                ("kT", C.c_double), ("time_interval", C.c_double),
                ("ewald_neut", C.c_double), ("n_path", C.c_int64),
                ("max_hops", C.c_int64)]


class TrajResultC(C.Structure):
    _fields_ = [("t", C.c_double), ("energy", C.c_double),
                ("hops", C.c_int64), ("start", C.c_int64),
                ("no_select", C.c_int32), ("done", C.c_int32),
                ("busy_s", C.c_double), ("simd", C.c_int32),
                ("pad0", C.c_int32)]



_P = C.c_void_p
_SIGNATURES = {
    "pycd_abi_version": (C.c_int, []),
    "pycd_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pycd_ctx_destroy": (None, [_P]),
    "pycd_last_error": (C.c_char_p, [_P]),
    "pycd_ctx_cu_count": (C.c_int, [_P]),
    "pycd_pair_table": (C.c_int, [_P, _P, C.c_int64, _P, _P, _P, C.c_double,
                                  C.c_double, C.c_double, C.c_double,
                                  C.c_int32, _P, C.POINTER(EwaldReport)]),
    "pycd_min_image_vectors": (C.c_int, [_P, _P, C.c_int64, _P, _P, _P]),
    "pycd_hop_neighbors": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int64, _P,
                                     C.c_int64, _P, _P, _P, _P, C.c_int32,
                                     C.c_int32, _P, _P, _P]),
    "pycd_kmc_create": (C.c_int, [_P, C.POINTER(KmcProblemC),
                                  C.POINTER(C.c_void_p)]),
    "pycd_kmc_destroy": (None, [_P]),
    "pycd_kmc_begin": (C.c_int, [_P, _P, C.c_double]),
    "pycd_kmc_set_sites": (C.c_int, [_P, _P, _P]),
    "pycd_kmc_table_rows": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "pycd_set_table_variant": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "pycd_table_variant": (C.c_int, [_P, C.c_int32]),
    "pycd_kmc_set_variant": (C.c_int, [_P, C.c_int32]),
    "pycd_kmc_variant": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "pycd_kmc_advance": (C.c_int, [_P, _P, C.c_int64, C.c_int64, _P, _P,
                                   C.POINTER(C.c_int64)]),
    "pycd_kmc_advance_static": (C.c_int, [_P, _P, C.c_int64, C.c_int64, _P,
                                          _P, C.POINTER(C.c_int64)]),
    "pycd_kmc_read": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "pycd_msd": (C.c_int, [_P, _P, C.c_int64, C.c_int64, C.c_int64,
                           C.c_int64, _P]),
    "pycd_batch_create": (C.c_int, [C.c_void_p, C.POINTER(FamilyC), C.c_int32,
                                    C.c_double, C.c_int32, C.c_int32,
                                    C.POINTER(C.c_void_p)]),
    "pycd_batch_destroy": (None, [C.c_void_p]),
    "pycd_batch_setup": (C.c_int, [C.c_void_p, C.POINTER(TrajC), C.c_int64,
                                   C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_double, C.c_int32]),
    "pycd_batch_run": (C.c_int, [C.c_void_p, C.c_int32,
                                 C.POINTER(C.c_double)]),
    "pycd_batch_set_handover": (C.c_int, [C.c_void_p, C.c_int32]),
    "pycd_batch_set_hop_differences": (C.c_int, [C.c_void_p, C.c_int32]),
    "pycd_batch_results": (C.c_int, [C.c_void_p, C.POINTER(TrajResultC),
                                     C.c_void_p]),
    "pycd_batch_rows": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "pycd_batch_msd": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pycd_batch_kernel_ms": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pycd_batch_handovers": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "pycd_batch_resident": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "pycd_batch_add_family": (C.c_int, [_P, C.POINTER(FamilyC),
                                        C.POINTER(C.c_int32)]),
    "pycd_batch_drop_family": (C.c_int, [_P, C.c_int32]),
    "pycd_batch_pool": (C.c_int, [_P, C.c_int64, C.c_int32, C.c_int32,
                                  C.c_int32, C.c_int64, C.c_int32]),
    "pycd_batch_admit": (C.c_int, [_P, C.c_int64, C.POINTER(TrajC), _P, _P,
                                   C.c_int64]),
    "pycd_batch_release": (C.c_int, [_P, C.c_int64]),
    "pycd_batch_step": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_double)]),
    "pycd_batch_slot_msd": (C.c_int, [_P, _P, C.c_int64, _P, C.c_double,
                                      _P]),
    "pycd_dev_alloc": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_void_p)]),
    "pycd_dev_free": (C.c_int, [_P, _P]),
    "pycd_memcpy": (C.c_int, [_P, _P, _P, C.c_int64]),
    "pycd_kmc_last_ms": (C.c_double, [_P]),
}


def exported_symbols():
    """Names the binding expects the library to export."""
    return sorted(_SIGNATURES)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeError(
                f"{LIB_PATH} is missing; build it with "
                f"`python -m pycd_amd.build` (there is no CPU fallback)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check(ctx, rc, what):
    if rc < 0:
        msg = lib().pycd_last_error(ctx)
        raise NativeError(f"{what} failed ({rc}): "
                          f"{msg.decode() if msg else '?'}")
    return rc


def context(device=None):
    """The process-wide context of `device` (default: PYCD_DEVICE or 0)."""
    if device is None:
        device = int(os.environ.get("PYCD_DEVICE", "0"))
    if device not in _ctx:
        h = C.c_void_p()
        rc = lib().pycd_ctx_create(device, C.byref(h))
        if rc != 0:
            raise NativeError(f"pycd_ctx_create({device}) failed ({rc}): no "
                              f"usable HIP device")
        _ctx[device] = h
    return _ctx[device]


def device_cus(device=None):
    """Compute units of the device behind `context(device)`."""
    return int(lib().pycd_ctx_cu_count(context(device)))


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


# ------------------------------------------------------------------ Ewald --
def pair_table(pos, cell, trans, pbc, alpha, r_cut, k_cut, eps_r,
               precision="f64", out=None):
    """N x N Ewald table on the GPU; returns (table, report dict)."""
    pos = _f64(pos)
    n = pos.shape[0]
    cell = _f64(cell, (9,))
    trans = _f64(trans, (9,))
    pbc = np.ascontiguousarray(pbc, dtype=np.int32).reshape(3)
    if out is None:
        out = np.empty((n, n), dtype=np.float64)
    rep = EwaldReport()
    ctx = context()
    rc = lib().pycd_pair_table(ctx, _ptr(pos), n, _ptr(cell), _ptr(trans),
                               _ptr(pbc), float(alpha), float(r_cut),
                               float(k_cut), float(eps_r), PREC[precision],
                               _ptr(out), C.byref(rep))
    _check(ctx, rc, "pycd_pair_table")
    report = {"n_kvec": rep.n_kvec, "n_real_pairs": rep.n_real_pairs,
              "k_max": list(rep.k_max), "precision": precision,
              "volume": rep.volume, "ms_sf": rep.ms_sf,
              "ms_real": rep.ms_real, "ms_recip": rep.ms_recip,
              "ms_total": rep.ms_total}
    return out, report


def set_table_variant(precision, variant):
    """Tile variant of the table kernels for `precision` in this process's
    context (0 = default; FP64: 1 = 64 x 64 tiles, 2 = 128 x 128; BF16x3:
    1 = K slices of 32, 2 = of 64, 3 = of 32 in two LDS stages, 4 = 256 x 256
    tiles in two LDS stages; FP32: 1).  Every variant gives the same table
    bit for bit."""
    ctx = context()
    _check(ctx, lib().pycd_set_table_variant(ctx, PREC[precision],
                                             int(variant)),
           "pycd_set_table_variant")


def table_variant(precision):
    """The table kernel variant set for `precision` (0 = default)."""
    return int(lib().pycd_table_variant(context(), PREC[precision]))


TABLE_VARIANTS = {"f64": (1, 2), "f32": (1,), "bf16x3": (1, 2, 3, 4)}


class DeviceArray:
    """A device buffer of the context's device (bytes + dtype + shape)."""

    def __init__(self, nbytes, dtype, shape):
        self.ctx = context()
        h = C.c_void_p()
        _check(self.ctx, lib().pycd_dev_alloc(self.ctx, int(nbytes),
                                              C.byref(h)), "pycd_dev_alloc")
        self.ptr = h
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape)

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, a.dtype, a.shape)
        _check(d.ctx, lib().pycd_memcpy(d.ctx, d.ptr, _ptr(a), a.nbytes),
               "pycd_memcpy")
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.ctx, lib().pycd_memcpy(self.ctx, _ptr(out), self.ptr,
                                           self.nbytes), "pycd_memcpy")
        return out

    def free(self):
        if self.ptr:
            lib().pycd_dev_free(self.ctx, self.ptr)
            self.ptr = None


# ------------------------------------------------------------- neighbours --
def min_image_vectors(pos, cell, pbc):
    pos = _f64(pos)
    n = pos.shape[0]
    out = np.empty((n, n, 3))
    cell = _f64(cell, (9,))
    pbc = np.ascontiguousarray(pbc, dtype=np.int32).reshape(3)
    ctx = context()
    _check(ctx, lib().pycd_min_image_vectors(ctx, _ptr(pos), n, _ptr(cell),
                                             _ptr(pbc), _ptr(out)),
           "pycd_min_image_vectors")
    return out


def hop_neighbors(pos, centers, targets, cell, pbc, lo, hi, max_per=64):
    """Per centre and shell: (positions into targets, vectors)."""
    pos = _f64(pos)
    centers = np.ascontiguousarray(centers, dtype=np.int64)
    targets = np.ascontiguousarray(targets, dtype=np.int64)
    lo = _f64(lo)
    hi = _f64(hi)
    ns = len(lo)
    cell = _f64(cell, (9,))
    pbc = np.ascontiguousarray(pbc, dtype=np.int32).reshape(3)
    ctx = context()
    while True:
        count = np.empty((len(centers), ns), dtype=np.int32)
        index = np.empty((len(centers), ns, max_per), dtype=np.int32)
        vec = np.empty((len(centers), ns, max_per, 3))
        rc = lib().pycd_hop_neighbors(ctx, _ptr(pos), len(pos), _ptr(centers),
                                      len(centers), _ptr(targets),
                                      len(targets), _ptr(cell), _ptr(pbc),
                                      _ptr(lo), _ptr(hi), ns, max_per,
                                      _ptr(count), _ptr(index), _ptr(vec))
        if rc < 0 and count.max() > max_per:
            max_per = int(count.max())
            continue
        _check(ctx, rc, "pycd_hop_neighbors")
        return count, index, vec


# -------------------------------------------------------------------- KMC --
class KmcProblem:
    """Host arrays of one trajectory family (atomic units)."""

    def __init__(self, table, lattice_q, nbr_count, nbr_site, nbr_vec,
                 nbr_lambda, nbr_vab, site_energy, carrier_q, field, nu, kT,
                 time_interval, n_path, cells=None):
        """`cells`: the supercell size of a periodic lattice (sites in
        PyCD's cell-major order), so that the hop loop may gather from the
        table rows of cell 0; None: gather from the full table."""
        self.table = _f64(table)
        n = self.table.shape[0]
        self.lattice_q = _f64(lattice_q)
        self.nbr_count = np.ascontiguousarray(nbr_count, dtype=np.int32)
        self.nbr_site = np.ascontiguousarray(nbr_site, dtype=np.int32)
        self.max_z = self.nbr_site.shape[1]
        self.nbr_vec = _f64(nbr_vec)
        self.nbr_lambda = _f64(nbr_lambda)
        self.nbr_vab = _f64(nbr_vab)
        self.site_energy = _f64(site_energy)
        self.carrier_q = _f64(carrier_q)
        self.field = _f64(field)
        self.nu, self.kT = float(nu), float(kT)
        self.time_interval = float(time_interval)
        self.n_path = int(n_path)
        counts = set(int(c) for c in self.nbr_count if c > 0)
        if len(counts) != 1:
            raise ValueError("every hopping site must have the same number "
                             f"of hop slots (found {sorted(counts)})")
        self.z = counts.pop()
        assert self.lattice_q.shape == (n,)
        self.cells = (0, 0, 0)
        self.n_unit = 0
        if cells is not None:
            cells = tuple(int(v) for v in cells)
            n_cells = cells[0] * cells[1] * cells[2]
            if n_cells > 0 and n % n_cells == 0:
                self.cells, self.n_unit = cells, n // n_cells


class KmcSession:
    """One device-resident KMC family; run() executes one trajectory."""

    def __init__(self, problem, chunk=1 << 16, max_chunk=1 << 21):
        self.pb = problem
        self.ctx = context()
        self.chunk0 = chunk
        self.max_chunk = max_chunk
        pb = problem
        c = KmcProblemC()
        c.n_sites = pb.table.shape[0]
        c.n_carriers = len(pb.carrier_q)
        c.z = pb.z
        c.max_z = pb.max_z
        c.table = pb.table.ctypes.data
        c.lattice_q = pb.lattice_q.ctypes.data
        c.nbr_site = pb.nbr_site.ctypes.data
        c.nbr_vec = pb.nbr_vec.ctypes.data
        c.nbr_lambda = pb.nbr_lambda.ctypes.data
        c.nbr_vab = pb.nbr_vab.ctypes.data
        c.site_energy = pb.site_energy.ctypes.data
        c.carrier_q = pb.carrier_q.ctypes.data
        for a in range(3):
            c.field[a] = float(pb.field[a])
        c.nu, c.kT = pb.nu, pb.kT
        c.time_interval = pb.time_interval
        c.n_path = pb.n_path
        for a in range(3):
            c.cells[a] = pb.cells[a]
        c.n_unit = pb.n_unit
        h = C.c_void_p()
        _check(self.ctx, lib().pycd_kmc_create(self.ctx, C.byref(c),
                                               C.byref(h)), "pycd_kmc_create")
        self.h = h

    def close(self):
        if self.h:
            lib().pycd_kmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def table_rows(self):
        """Table rows the hop loop gathers from (n_unit: cell-0 rows of a
        periodic supercell; n_sites: the full table)."""
        n = C.c_int64(0)
        _check(self.ctx, lib().pycd_kmc_table_rows(self.h, C.byref(n)),
               "pycd_kmc_table_rows")
        return n.value

    def set_variant(self, name):
        """Kernel of the hop loop: 'auto' (default), 'workgroup' or 'wave'
        (the one-wavefront kernel with its LDS row cache; refused where it
        does not apply)."""
        _check(self.ctx, lib().pycd_kmc_set_variant(self.h, KMC_VARIANT[name]),
               "pycd_kmc_set_variant")

    def variant(self):
        """Kernel the next advance runs: 'wave' or 'workgroup'."""
        v = C.c_int32(0)
        _check(self.ctx, lib().pycd_kmc_variant(self.h, C.byref(v)),
               "pycd_kmc_variant")
        return {1: "workgroup", 2: "wave"}[v.value]

    def set_sites(self, lattice_q, site_energy):
        """Lattice charges and site energies of the next trajectories."""
        n = self.pb.table.shape[0]
        q = _f64(lattice_q)
        e = _f64(site_energy)
        if q.shape != (n,) or e.shape != (n,):
            raise ValueError(f"lattice_q and site_energy must have {n} "
                             f"entries")
        _check(self.ctx, lib().pycd_kmc_set_sites(self.h, _ptr(q), _ptr(e)),
               "pycd_kmc_set_sites")

    def begin(self, occupancy, ewald_neut=0.0):
        occ = np.ascontiguousarray(occupancy, dtype=np.int64)
        _check(self.ctx, lib().pycd_kmc_begin(self.h, _ptr(occ),
                                              float(ewald_neut)),
               "pycd_kmc_begin")

    def advance(self, uniforms, max_hops=None, record=False):
        """Consume pairs of `uniforms`; returns (rc, times, procs)."""
        u = _f64(uniforms)
        n_pairs = u.size // 2
        times = np.empty(max(n_pairs, 1), dtype=np.float64)
        procs = np.empty(max(n_pairs, 1), dtype=np.int32) if record else None
        made = C.c_int64(0)
        rc = lib().pycd_kmc_advance(self.h, _ptr(u), n_pairs,
                                    int(max_hops or 0), _ptr(times),
                                    _ptr(procs) if record else None,
                                    C.byref(made))
        _check(self.ctx, rc, "pycd_kmc_advance")
        k = made.value
        return rc, times[:k], (procs[:k] if record else None)

    def advance_device(self, dev_uniforms, n_pairs, max_hops=None):
        """Advance from uniforms already in device memory; returns the
        kernel's device time in ms."""
        made = C.c_int64(0)
        rc = lib().pycd_kmc_advance(self.h, dev_uniforms.ptr, int(n_pairs),
                                    int(max_hops or 0), None, None,
                                    C.byref(made))
        _check(self.ctx, rc, "pycd_kmc_advance")
        self.last_made = made.value
        return lib().pycd_kmc_last_ms(self.h)

    def advance_static(self, dev_uniforms, n_pairs, max_hops=None,
                       dev_times=None, dev_procs=None, offset_pairs=0):
        """pycd_kmc_advance_static: device buffers only, nothing allocated;
        uniforms from pair `offset_pairs` of `dev_uniforms`.  Returns
        (rc, hops made)."""
        made = C.c_int64(0)
        u = C.c_void_p(dev_uniforms.ptr.value + 16 * int(offset_pairs))
        rc = lib().pycd_kmc_advance_static(
            self.h, u, int(n_pairs), int(max_hops or 0),
            dev_times.ptr if dev_times is not None else None,
            dev_procs.ptr if dev_procs is not None else None, C.byref(made))
        _check(self.ctx, rc, "pycd_kmc_advance_static")
        return rc, made.value

    def read(self):
        pb = self.pb
        nc = len(pb.carrier_q)
        P = pb.n_path
        rows = np.empty((P, 3 * nc))
        energy = np.empty(P)
        delg = np.empty(P)
        drift = np.empty((nc, 3))
        occ = np.empty(nc, dtype=np.int64)
        sc = np.empty(4)
        _check(self.ctx, lib().pycd_kmc_read(self.h, _ptr(rows), _ptr(energy),
                                             _ptr(delg), _ptr(drift),
                                             _ptr(occ), _ptr(sc)),
               "pycd_kmc_read")
        return {"unwrapped": rows, "energy": energy, "delg_0": delg,
                "drift": drift, "occupancy": [int(v) for v in occ],
                "t": sc[0], "energy_now": sc[1], "start": int(sc[2]),
                "hops": int(sc[3])}

    def run(self, occupancy, stream, ewald_neut=0.0, max_hops=None,
            record=False):
        """Whole trajectory from `occupancy`, uniforms from `stream.take`."""
        self.begin(occupancy, ewald_neut)
        chunk = self.chunk0
        times, procs = [np.zeros(1)], []
        while True:
            if max_hops is not None:
                done = sum(len(t) for t in times) - 1
                chunk = max(1, min(chunk, int(max_hops) - done))
            u = stream.take(2 * chunk)
            rc, t, p = self.advance(u, max_hops=max_hops, record=record)
            times.append(t)
            if record:
                procs.append(p)
            if rc == OK:
                break
            chunk = min(chunk * 2, self.max_chunk)
        out = self.read()
        out["time_data"] = np.concatenate(times)
        out["procs"] = (np.concatenate(procs).astype(np.int64) if record
                        else None)
        return out


# -------------------------------------------------------------------- MSD --
def msd(pos, n_lags):
    """sd[t, tau, c] for pos of shape (n_traj, P, n_c, 3)."""
    pos = _f64(pos)
    n_traj, P, nc, three = pos.shape
    assert three == 3
    sd = np.empty((n_traj, int(n_lags), nc))
    ctx = context()
    _check(ctx, lib().pycd_msd(ctx, _ptr(pos), n_traj, P, nc, int(n_lags),
                               _ptr(sd)), "pycd_msd")
    return sd

"""Synthetic KMC workloads: the lattices and run parameters of the benchmark
configurations, built through the public API (Material/Neighbors/System/Run).

  corundum  Fe-only corundum cell: hexagonal a = 5.038 A, c = 13.772 A,
            12 sites at Wyckoff 12c (0,0,+-z), (0,0,1/2+-z), z = 0.3553,
            plus the R-centring translations (2/3,1/3,1/3), (1/3,2/3,2/3).
            Hops: basal shell 2.971 A (3 neighbours), c shell 2.901 A
            (1 neighbour), tolerance 0.002 A.  lambda = 1.0 eV,
            V_ab = 0.2 / 0.1 eV, nu = 1.85e13 /s, Fe charge +3, electron -1,
            eps_r = 1, alpha = 0.18 /A, r_cut = 12 A, k_cut = 2 pi m / L_a.
  monoclinic  a = 7.25, b = 11.70, c = 5.09 A, beta = 134.2 deg, 4 sites at
            (0,1/4,1/8), (0,3/4,7/8), (1/2,3/4,1/8), (1/2,1/4,7/8);
            nearest-neighbour hops up to 4.2 A, lambda = 0.8 eV,
            V_ab = 0.1 eV, nu = 1e13 /s.

Values not fixed by the configuration text (nu, lattice charge) are taken
from the reference's own example inputs (examples/Hematite and
examples/BVO sys_config.yml).
"""
import math
import tempfile
from pathlib import Path

import numpy as np

from . import constants
from .core import Material, Neighbors, System, Run, ReturnValues

_TMP = None


def _tmpdir():
    global _TMP
    if _TMP is None:
        _TMP = Path(tempfile.mkdtemp(prefix="pycd_amd_"))
    return _TMP


def _write_poscar(name, lattice, symbol, frac):
    lines = [name, "1.0"]
    for row in lattice:
        lines.append("  " + "  ".join(f"{v:.10f}" for v in row))
    lines.append(f"   {symbol}")
    lines.append(f"   {len(frac)}")
    lines.append("Direct")
    for f in frac:
        lines.append("  " + "  ".join(f"{v:.12f}" for v in f))
    path = _tmpdir() / f"{name}.POSCAR"
    path.write_text("\n".join(lines) + "\n")
    return path


def corundum_fractional(zpar=0.3553):
    base = [(0.0, 0.0, zpar), (0.0, 0.0, -zpar), (0.0, 0.0, 0.5 + zpar),
            (0.0, 0.0, 0.5 - zpar)]
    shifts = [(0.0, 0.0, 0.0), (2 / 3, 1 / 3, 1 / 3), (1 / 3, 2 / 3, 2 / 3)]
    out = []
    for s in shifts:
        for b in base:
            out.append([(b[k] + s[k]) % 1.0 for k in range(3)])
    return np.array(out)


def corundum_params(m_kcut=6, size=3, lam=1.0, vab=(0.2, 0.1)):
    a, c = 5.038, 13.772
    lattice = np.array([[a, 0.0, 0.0], [-a / 2, a * math.sqrt(3) / 2, 0.0],
                        [0.0, 0.0, c]])
    poscar = _write_poscar("corundum_fe", lattice, "Fe", corundum_fractional())
    params = {
        "name": "Fe_corundum",
        "species_types": ["electron"],
        "species_charge_list": {"full": [-1]},
        "species_to_element_type_map": {"electron": ["Fe"]},
        "charge_types": {"full": {"Fe": 3.0}},
        "neighbor_cutoff_dist": {"Fe:Fe": [[2.971, 2.901]]},
        "neighbor_cutoff_dist_tol": {"Fe:Fe": [[0.002, 0.002]]},
        "vn": 1.85e13,
        "lambda_values": {"Fe:Fe": [[lam, lam]]},
        "v_ab": {"Fe:Fe": [[vab[0], vab[1]]]},
        "class_list": {"Fe": [1] * 12},
        "element_type_delimiter": ":",
        "dielectric_constant": 1.0,
        "alpha": 0.18,
        "r_cut": 12.0,
        "k_cut": 2 * math.pi * m_kcut / (size * a),
        "precision_parameters": {"err_tol": 2.585e-03},
        "input_coord_file_location": poscar,
    }
    return params


def monoclinic_params(lam=0.8, vab=0.1, shells=None):
    a, b, c, beta = 7.25, 11.70, 5.09, math.radians(134.2)
    lattice = np.array([[a, 0.0, 0.0], [0.0, b, 0.0],
                        [c * math.cos(beta), 0.0, c * math.sin(beta)]])
    frac = np.array([[0.0, 0.25, 0.125], [0.0, 0.75, 0.875],
                     [0.5, 0.75, 0.125], [0.5, 0.25, 0.875]])
    poscar = _write_poscar("monoclinic_v", lattice, "V", frac)
    if shells is None:
        shells = monoclinic_shells(lattice, frac)
    params = {
        "name": "monoclinic_V",
        "species_types": ["electron"],
        "species_charge_list": {"full": [-1]},
        "species_to_element_type_map": {"electron": ["V"]},
        "charge_types": {"full": {"V": 5.0}},
        "neighbor_cutoff_dist": {"V:V": [list(shells)]},
        "neighbor_cutoff_dist_tol": {"V:V": [[0.002] * len(shells)]},
        "vn": 1.00e13,
        "lambda_values": {"V:V": [[lam] * len(shells)]},
        "v_ab": {"V:V": [[vab] * len(shells)]},
        "class_list": {"V": [1] * 4},
        "element_type_delimiter": ":",
        "dielectric_constant": 1.0,
        "alpha": 0.18,
        "r_cut": 12.0,
        "k_cut": 2 * math.pi * 6 / (4 * 7.25),
        "precision_parameters": {"err_tol": 2.585e-03},
        "input_coord_file_location": poscar,
    }
    return params


def monoclinic_shells(lattice, frac, r_max=4.2):
    """Distinct site-site distances (A, rounded to 1e-4) up to r_max."""
    cart = frac @ lattice
    dists = set()
    for i in range(len(cart)):
        for j in range(len(cart)):
            for n0 in range(-2, 3):
                for n1 in range(-2, 3):
                    for n2 in range(-2, 3):
                        d = cart[j] - cart[i] + np.array([n0, n1, n2]) @ lattice
                        r = float(np.linalg.norm(d))
                        if 1e-6 < r <= r_max:
                            dists.add(round(r, 4))
    return sorted(dists)


class Workload:
    """Material + Neighbors + System (+ table) + Run for one configuration."""

    def __init__(self, params, size, temp=300.0, n_carriers=1, field_vpm=0.0,
                 field_dir=(0, 0, 1), field_ld=0, t_final=1e-7,
                 time_interval=1e-12, seed=2, table_precision="f64",
                 hop_lists=None):
        """hop_lists(material, neighbors) may supply the hop neighbour list
        (default: built on the GPU)."""
        self.params = params
        self.size = np.array([size] * 3 if np.isscalar(size) else size)
        self.material = Material(ReturnValues(**params))
        self.neighbors = Neighbors(self.material, self.size,
                                   np.array([1, 1, 1]))
        hop = hop_lists(self.material, self.neighbors) if hop_lists else None
        self.system = System(self.material, self.neighbors, hop, None,
                             params["alpha"], params["r_cut"], params["k_cut"],
                             params["precision_parameters"], [], [],
                             table_precision=table_precision)
        self.temp = temp
        self.n_carriers = n_carriers
        self.field = {"electric": {
            "active": 1 if field_vpm else 0, "dir": list(field_dir),
            "ld": field_ld,
            "mag": field_vpm * constants.VPM2AUFIELD}}
        self.t_final = t_final
        self.time_interval = time_interval
        self.seed = seed
        self.table = None

    @property
    def n_sites(self):
        return self.neighbors.num_system_elements

    def build_table(self, dst_path=None):
        self.table, _ = self.system.get_precomputed_array(dst_path, 0, 0)
        return self.table

    def make_run(self, table=None, n_traj=1):
        table = self.table if table is None else table
        return Run(self.system, table, self.temp, "full", "full", n_traj,
                   self.t_final, self.time_interval,
                   np.array([self.n_carriers]), {"electron": []},
                   {"class_index": {}, "doping": {}}, self.field,
                   {"num_dopants": [0]})


def corundum(size, **kw):
    m = kw.pop("m_kcut", 6)
    return Workload(corundum_params(m_kcut=m, size=size), size, **kw)


def monoclinic(size, **kw):
    return Workload(monoclinic_params(), size, **kw)


def batched_config(n_traj=4096, hops=100_000, seed=2, param_seed=1234,
                   edges=(4, 5, 6, 7, 8), nc_range=(1, 64),
                   field_max_vpm=1e7, t_range=(250.0, 400.0),
                   path_rows=1000, msd_lags=500, pilot_hops=3000):
    """The batched workload: n_traj trajectories of uneven size and load.

    Edge n ~ U{edges}, carriers ~ U{nc_range}, |E| ~ U(0, field_max) V/m
    along a random lattice direction (`ld` = 1: fractional direction turned
    Cartesian by the lattice), T ~ U(t_range) K, `hops` hops each.  The
    trajectory parameters come from a NumPy generator seeded `param_seed`;
    the KMC streams follow the reference's seeding protocol from `seed`.
    The path-step width is chosen per trajectory so that `hops` hops span
    about `path_rows` rows (2 x path_rows + 1 rows are allocated), from the
    clock of a `pilot_hops`-hop run of the same trajectory on the GPU (the
    zero-interaction rate estimate alone is several times too fast for
    many carriers, which would end trajectories early on a full path);
    `pilot_hops=0` keeps the estimate.  Every trajectory then runs its
    `hops` hops (the configuration's step count) unless the path fills.
    """
    from . import batch as B
    from .rng import seeded_starts
    gen = np.random.default_rng(param_seed)
    fams, works = [], []
    for n in edges:
        w = corundum(n, m_kcut=6)
        table = w.build_table()
        run = w.make_run(table)
        fams.append(B.Family(run, table))
        works.append(w)
    sizes = gen.integers(0, len(edges), size=n_traj)
    ncs = gen.integers(nc_range[0], nc_range[1] + 1, size=n_traj)
    mags = gen.uniform(0.0, field_max_vpm, size=n_traj)
    temps = gen.uniform(t_range[0], t_range[1], size=n_traj)
    dirs = gen.integers(-1, 2, size=(n_traj, 3))
    for k in range(n_traj):
        while not np.any(dirs[k]):
            dirs[k] = gen.integers(-1, 2, size=3)
    occs, words = seeded_starts(seed, n_traj,
                                [fams[s].hop_sites for s in sizes], ncs)
    trajs = []
    for k in range(n_traj):
        fam = fams[sizes[k]]
        lat = fam.run.material.lattice_matrix
        d = dirs[k] @ lat
        field = mags[k] * constants.VPM2AUFIELD * d / np.linalg.norm(d)
        kT = temps[k] * constants.K2AUTEMP
        rate = ncs[k] * B.expected_rate_per_carrier(fam, kT, field)
        dt = hops / rate / path_rows
        trajs.append(dict(family=int(sizes[k]), n_carriers=int(ncs[k]),
                          carrier_q=-1.0, field=field, kT=kT,
                          time_interval=dt, ewald_neut=0.0,
                          n_path=2 * path_rows + 1, max_hops=hops))
    if pilot_hops > 0:
        bt = B.KmcBatch(fams, table_precision="f64", rate_precision="f32")
        try:
            pilot = [dict(t, max_hops=pilot_hops) for t in trajs]
            bt.setup(pilot, occs, words, waves_per_cu=8)
            bt.run()
            res = bt.results()
        finally:
            bt.close()
        for k, t in enumerate(trajs):
            h, tt = float(res["hops"][k]), float(res["t"][k])
            if h > 0 and np.isfinite(tt) and tt > 0:
                t["time_interval"] = hops * (tt / h) / path_rows
    return fams, works, trajs, occs, words, [msd_lags] * n_traj

"""Builders shared by the tests: the reference's two shipped examples."""
import types

import numpy as np

from pycd_amd.core import Material, Neighbors, System, Run, ReturnValues
from tests.conftest import golden_dir, load_configs


def build_example(material):
    cfg, sim = load_configs(material)
    d = golden_dir(material)
    mat = Material(ReturnValues(**cfg))
    nb = Neighbors(mat, np.array(sim["system_size"]), np.array(sim["pbc"]))
    table = np.load(d / "precomputed_array.npy")
    pairwise = np.load(d / "pairwise_min_image_vector_data.npy")
    from oracle.neighbors import hop_neighbor_list
    system = System(mat, nb, hop_neighbor_list(mat, nb), pairwise,
                    cfg["alpha"], cfg["r_cut"],
                    cfg["k_cut"], cfg["precision_parameters"], [], [])
    run = Run(system, table, sim["temp"], sim["ion_charge_type"],
              sim["species_charge_type"], sim["n_traj"], sim["t_final"],
              sim["time_interval"], np.array(sim["species_count"]),
              sim["initial_occupancy"], sim["relative_energies"],
              sim["external_field"], sim["doping"])
    # Site order of the shipped table: the reference's z-sort broke ties of
    # equal-z sites differently from this NumPy, so positions are rebuilt
    # from the shipped minimum-image vectors of site 0 (a pure function of
    # the shipped data; site 0 is a hopping site with a unique z).
    shipped_pos = nb.bulk_sites.cell_coordinates[0] + pairwise[0]
    return types.SimpleNamespace(
        name=material, cfg=cfg, sim=sim, material=mat, neighbors=nb,
        system=system, run=run, table=table, pairwise=pairwise,
        shipped_pos=shipped_pos, dir=d,
        time_data=np.load(d / "time_data.npy"),
        unwrapped=np.load(d / "unwrapped_traj.npy"),
        rnd_dump=(d / "initial_rnd_state.dump").read_bytes())


def oracle_problem(run, table=None, field=None):
    """oracle.kmc.Problem from a pycd_amd Run (same host tables)."""
    from oracle import kmc
    return kmc.Problem(
        run.precomputed_array if table is None else table,
        run.base_charge_config()[:, 0], run.nbr_count, run.nbr_site,
        run.nbr_vec, run.nbr_lambda, run.nbr_vab,
        run.system_relative_energies, run.species_charge_list,
        (run.electric_field if run.electric_field_active else np.zeros(3))
        if field is None else field,
        run.material.vn, run.temp, run.time_interval,
        int(run.t_final / run.time_interval) + 1)


def cell0_table(table, cells, n_unit):
    """The N x N table a cell-periodic family reads: entry (a, b) is the
    row of a's sublattice in cell 0 at b moved by -cell(a), with sites
    numbered ((cx ny + cy) nz + cz) n_unit + sub.  It equals `table` to
    rounding (the family's set-up checks that); running the oracle on it
    compares the batched kernel with the reference's loop on the very
    entries the kernel reads."""
    nx, ny, nz = (int(v) for v in cells)
    idx = np.arange(nx * ny * nz * n_unit)
    sub = idx % n_unit
    c = idx // n_unit
    cz, cy, cx = c % nz, (c // nz) % ny, c // (nz * ny)
    dx = (cx[None, :] - cx[:, None]) % nx
    dy = (cy[None, :] - cy[:, None]) % ny
    dz = (cz[None, :] - cz[:, None]) % nz
    col = ((dx * ny + dy) * nz + dz) * n_unit + sub[None, :]
    return np.ascontiguousarray(table[sub[:, None], col])


DOPING_SHIFTS = [0.25, -0.05, 0.01]  # eV at hop shells 0, 1, 2 of a dopant


def doped_run(ex, num, kind="random", sep=1, initiation="yes", manual=None,
              n_traj=3, charge=4.0):
    """The example's Run with Ti substituting Fe (charge `charge`) at `num`
    sites, site energies shifted by DOPING_SHIFTS around each dopant."""
    sim = ex.sim
    spec = {"doping_element_map": ["Fe:Ti"],
            "charge": {"full": {"Ti": charge}},
            "num_dopants": [num], "insertion_type": [kind],
            "min_shell_separation": [sep], "max_attempts": 20,
            "site_charge_initiation": [initiation],
            "dopant_site_indices": [manual or []]}
    energies = {"class_index": sim["relative_energies"]["class_index"],
                "doping": {"Fe": [DOPING_SHIFTS]}}
    return Run(ex.system, ex.table, sim["temp"], sim["ion_charge_type"],
               sim["species_charge_type"], n_traj, sim["t_final"],
               sim["time_interval"], np.array(sim["species_count"]),
               sim["initial_occupancy"], energies, sim["external_field"],
               spec)

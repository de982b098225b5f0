"""Doped trajectories on the HIP plain path.

A doped trajectory is an undoped one with other lattice charges and site
energies (and another start), so the oracle checks it unchanged, given the
doped arrays: the same events hop by hop, clock within 1e-12 relative.
The placement itself is checked on the CPU (tests/test_doping.py).
"""
import pickle
import random

import numpy as np
import pytest

from oracle import kmc
from pycd_amd import _native, rng
from tests.kmc_helpers import native_problem, seeded_start
from tests.systems import doped_run

pytestmark = pytest.mark.gpu

TIME_TOL = 1e-12  # relative, per hop


def _problem_with(run, table, lattice_q, site_energy):
    pb = native_problem(run, table)
    pb.lattice_q = np.ascontiguousarray(lattice_q, dtype=np.float64)
    pb.site_energy = np.ascontiguousarray(site_energy, dtype=np.float64)
    return pb


def _same(a, b):
    for key in ("time_data", "unwrapped", "energy", "delg_0", "procs"):
        assert np.array_equal(a[key], b[key]), key
    assert a["energy_now"] == b["energy_now"]


def test_set_sites_equals_a_session_built_with_those_sites(hematite,
                                                           tmp_path):
    run = doped_run(hematite, 3, charge=3.2)
    run.preproduction(tmp_path, 5)
    energy, q, _ = run.doped_inputs(tmp_path / "traj1")
    gen, occ = seeded_start(hematite.run, 2, 1)
    state = gen.getstate()
    n = 4000

    def go(sess):
        return sess.run(occ, rng.UniformStream(state), max_hops=n,
                        record=True)

    plain = _native.KmcSession(native_problem(hematite.run, hematite.table))
    fresh = _native.KmcSession(_problem_with(hematite.run, hematite.table,
                                             q, energy))
    try:
        r0 = go(plain)
        plain.set_sites(q, energy)
        r1 = go(plain)
        r2 = go(fresh)
        _same(r1, r2)
        assert r1["energy_now"] != r0["energy_now"]
        # back to the undoped sites: the undoped run again, bit for bit
        plain.set_sites(hematite.run.base_charge_config()[:, 0],
                        hematite.run.system_relative_energies)
        _same(go(plain), r0)
        with pytest.raises(ValueError):
            plain.set_sites(q[:-1], energy)
    finally:
        plain.close()
        fresh.close()


def test_doped_do_kmc_steps_vs_oracle(hematite, tmp_path):
    # Ti at +3.2 against Fe +3: a shallow donor.  At +4 the electron, placed
    # on a dopant, waits past t_final on its first hop (the oracle agrees).
    run = doped_run(hematite, 3, charge=3.2)
    n = 5000
    run.do_kmc_steps(tmp_path, hematite.sim["output_data"], 7, "serial",
                     max_hops=n, record=True)
    system_charge = np.dot(run.species_count,
                           run.material.species_charge_list[
                               run.species_charge_type])
    ewald_neut = -(np.pi * system_charge ** 2
                   / (2 * run.system.system_volume * run.system.alpha))
    for k in range(run.n_traj):
        d = tmp_path / f"traj{k + 1}"
        energy, q, sites = run.doped_inputs(d)
        with open(d / "initial_rnd_state.dump", "rb") as fh:
            random.setstate(pickle.load(fh))
        occ = run.generate_initial_occupancy(sites)
        assert occ[0] in sites["Ti"]   # site_charge_initiation 'yes'
        pb = kmc.Problem(run.precomputed_array, q, run.nbr_count,
                         run.nbr_site, run.nbr_vec, run.nbr_lambda,
                         run.nbr_vab, energy, run.species_charge_list,
                         np.zeros(3), run.material.vn, run.temp,
                         run.time_interval,
                         int(run.t_final / run.time_interval) + 1)
        ora = kmc.run(pb, occ, random.random, max_hops=n, record=True,
                      ewald_neut=ewald_neut)
        gpu = run.last_records[k]
        print(f"traj{k + 1}: dopants {sites['Ti']}, {gpu['hops']} hops")
        assert gpu["hops"] == ora["hops"] > 100
        assert np.array_equal(gpu["procs"], ora["procs"])
        rel = np.abs(gpu["time_data"][1:] - ora["time_data"][1:]) / \
            ora["time_data"][1:]
        assert rel.max() < TIME_TOL
        assert np.max(np.abs(gpu["unwrapped"] - ora["unwrapped"])) < 1e-8
        scale = max(1.0, abs(ora["energy_now"]))
        assert abs(gpu["energy_now"] - ora["energy_now"]) < 1e-10 * scale
        # the written path is the recorded one
        ut = np.load(d / "unwrapped_traj.npy")
        assert np.array_equal(ut, gpu["unwrapped"])
        # occupancy before the first hop and after every hop, as the loop
        # itself held it
        occ_rows = np.load(d / "occupancy.npy")
        assert occ_rows.shape == (n + 1, len(occ))
        assert np.array_equal(occ_rows, ora["occ_rows"])

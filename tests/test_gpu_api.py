"""Drop-in use of Run.do_kmc_steps and Analysis.compute_msd."""
import pickle

import numpy as np
import pytest

from oracle import msd as omsd, seeding
from pycd_amd import Analysis, constants, workloads

pytestmark = pytest.mark.gpu

# dG0 is a difference of potentials of order 1 Ha: the GPU forms q.A[x, :]
# as phi_L + psi instead of the reference's N-long dot, which rounds
# differently at the 1e-15 Ha level whatever the size of dG0 itself
DG0_TOL = 1e-13  # Ha, absolute


def test_do_kmc_steps_reproduces_shipped_files(tmp_path, hematite):
    ex = hematite
    ex.run.do_kmc_steps(tmp_path, ex.sim["output_data"], ex.sim["random_seed"],
                        ex.sim["compute_mode"])
    traj = tmp_path / "traj1"
    # preproduction wrote the seeded state with this Python's default pickle
    # protocol (the shipped dump is protocol 3; its state is pinned by
    # test_seeding_protocol_bytes)
    state = seeding.trajectory_states(ex.sim["random_seed"], 1)[0]
    assert (traj / "initial_rnd_state.dump").read_bytes() == pickle.dumps(state)
    td = np.load(traj / "time_data.npy")
    assert len(td) == len(ex.time_data)
    assert np.max(np.abs(td[1:] - ex.time_data[1:]) / ex.time_data[1:]) < 1e-12
    ut = np.load(traj / "unwrapped_traj.npy")
    assert np.max(np.abs(ut - ex.unwrapped)) < 1e-9
    assert (tmp_path / "Run.log").exists()

    sim = ex.sim
    an = Analysis(ex.material, sim["n_dim"], np.array(sim["species_count"]),
                  sim["n_traj"], sim["t_final"], sim["time_interval"],
                  sim["msd_t_final"], sim["trim_length"], sim["temp"],
                  sim["repr_time"], sim["repr_dist"])
    out = an.compute_msd(tmp_path, sim["output_data"])
    P = an.num_path_steps_per_traj
    pos = ut.reshape(P, 1, 3) * (1 / constants.ANG2BOHR)
    ref = omsd.squared_displacement(pos, 1, P, an.num_msd_steps_per_traj)
    assert np.allclose(out.msd_data[:, 1], ref[0, :, 0], rtol=1e-12, atol=0)
    assert out.msd_data.shape == (501, 2)
    log = (tmp_path / f"MSD_Analysis_{out.file_name}.log").read_text()
    assert "electron diffusivity" in log


def test_compute_msd_three_trajectories_two_carriers(tmp_path, hematite):
    """Run.do_kmc_steps for three trajectories of two electrons, then
    Analysis.compute_msd: the saved MSD table, its SEM, the per-trajectory
    slopes and the logged D/kBT lines equal the oracle's restatement of
    PyCD/core.py:2974-3071 applied to the written trajectories."""
    from pycd_amd import Run
    ex = hematite
    sim = ex.sim
    run = Run(ex.system, ex.table, sim["temp"], sim["ion_charge_type"],
              sim["species_charge_type"], 3, sim["t_final"],
              sim["time_interval"], np.array([2, 0]),
              sim["initial_occupancy"], sim["relative_energies"],
              sim["external_field"], sim["doping"])
    run.do_kmc_steps(tmp_path, sim["output_data"], 5, sim["compute_mode"])
    an = Analysis(ex.material, sim["n_dim"], np.array([2, 0]), 3,
                  sim["t_final"], sim["time_interval"], sim["msd_t_final"],
                  sim["trim_length"], sim["temp"], sim["repr_time"],
                  sim["repr_dist"])
    out = an.compute_msd(tmp_path, sim["output_data"])
    P = an.num_path_steps_per_traj
    pos = np.vstack([np.load(tmp_path / f"traj{k + 1}" / "unwrapped_traj.npy")
                     for k in range(3)])
    pos = pos[:3 * P].reshape(3 * P, 2, 3) * (1 / constants.ANG2BOHR)
    sd = omsd.squared_displacement(pos, 3, P, an.num_msd_steps_per_traj)
    ref = omsd.reduce(sd, [2, 0], ex.material.species_types,
                      an.time_interval, an.time_conversion,
                      sim["trim_length"], sim["n_dim"], sim["temp"])
    assert np.allclose(out.msd_data, ref["msd_data"], rtol=1e-12, atol=0)
    assert np.allclose(out.sem_data, ref["sem_data"], rtol=1e-10, atol=0)
    assert np.allclose(an.slope_data, ref["slope_data"], rtol=1e-10, atol=0)
    assert out.species_types == ["electron"]
    log = (tmp_path / f"MSD_Analysis_{out.file_name}.log").read_text()
    for line in omsd.report_lines(ref["species_types"], ref["diffusivity"],
                                  ref["diffusivity_sem"]):
        assert line in log
    saved = np.load(tmp_path / f"MSD_Data_{out.file_name}.npy")
    assert np.array_equal(saved, out.msd_data)


def test_field_run_files_vs_oracle(tmp_path):
    """do_kmc_steps with a field along c, two trajectories of two electrons,
    energy and dG0 outputs on: time_data, unwrapped_traj, energy_traj,
    delG0_traj, drift_mobility.dat and the logged mobility equal the
    oracle's on the same seeds (PyCD/core.py:2647-2917, 2052-2082)."""
    import random
    import re
    from oracle import kmc
    from tests.systems import oracle_problem
    w = workloads.corundum(3, n_carriers=2, field_vpm=1e7, t_final=3e-10)
    table = w.build_table()
    run = w.make_run(table, n_traj=2)
    out_spec = {
        "unwrapped_traj": {"file_name": "unwrapped_traj.npy", "write": 1,
                           "write_every_step": 0},
        "time": {"file_name": "time_data.npy", "write": 1},
        "energy": {"file_name": "energy_traj.npy", "write": 1},
        "delg_0": {"file_name": "delG0_traj.npy", "write": 1},
        "wrapped_traj": {"file_name": "wrapped_traj.npy", "write": 0},
        "potential": {"file_name": "potential_traj.npy", "write": 0}}
    run.do_kmc_steps(tmp_path, out_spec, 3, "serial")
    charge = np.dot(run.species_count, run.material.species_charge_list[
        run.species_charge_type])
    neut = -(np.pi * charge ** 2
             / (2 * run.system.system_volume * run.system.alpha))
    drifts = []
    for k, state in enumerate(seeding.trajectory_states(3, 2)):
        gen = random.Random()
        gen.setstate(state)
        occ = seeding.initial_occupancy(gen, run.hop_site_lists["electron"],
                                        2)
        ora = kmc.run(oracle_problem(run, table=table), occ, gen.random,
                      ewald_neut=neut)
        d = tmp_path / f"traj{k + 1}"
        td = np.load(d / "time_data.npy")
        assert len(td) == len(ora["time_data"]) > 1000
        assert np.max(np.abs(td[1:] - ora["time_data"][1:])
                      / ora["time_data"][1:]) < 1e-12
        assert np.max(np.abs(np.load(d / "unwrapped_traj.npy")
                             - ora["unwrapped"])) < 1e-8
        e = np.load(d / "energy_traj.npy")
        assert np.max(np.abs(e - ora["energy"])) < 1e-10 * max(
            1.0, np.max(np.abs(ora["energy"])))
        g = np.load(d / "delG0_traj.npy")
        assert np.max(np.abs(g - ora["delg_0"])) < DG0_TOL
        drifts.append(ora["drift"])
    mob, per = kmc.drift_mobility(np.asarray(drifts), run.electric_field,
                                  run.electric_field_mag, run.species_count,
                                  run.material.species_types, 2)
    got = np.loadtxt(tmp_path / "drift_mobility.dat")
    assert np.allclose(got, mob, rtol=1e-9, atol=0)
    log = (tmp_path / "Run.log").read_text()
    vals = [float(v) for v in re.findall(
        r"drift mobility is: ([-+0-9.e]+) cm2/V\.s\.", log)]
    assert len(vals) == 2
    assert vals[0] == pytest.approx(per[0][1], rel=2e-3)
    assert vals[1] == pytest.approx(per[0][2], rel=2e-3)

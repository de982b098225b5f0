"""Batched FP32-rate trajectories reproduce the diffusivity within 2 sigma.

One electron on the 4x4x4 Fe corundum supercell (768 sites), zero field,
300 K: every dG0 is zero, so the exact D is known (oracle.diffusion).  The
batched kernel's MSD-derived D/kBT (Analysis.compute_msd procedure: MSD per
trajectory, least-squares slope over the trimmed lag window, mean and
standard error over trajectories) must lie within 2 SEM of it, and within
2 combined SEM of the FP64 oracle run on matched seeds.
"""
import random

import numpy as np
import pytest

from oracle import kmc, seeding
from oracle import msd as omsd
from oracle.diffusion import exact_diffusivity, over_kbt
from pycd_amd import batch as B, constants, rng, workloads
from pycd_amd.core import least_squares_slope
from tests.systems import oracle_problem

pytestmark = pytest.mark.gpu

HOPS, ROWS, LAGS, TRIM = 20000, 1000, 500, 50


def _d_over_kbt(curves, dt_au, temp):
    """Per-trajectory D/kBT from carrier-averaged MSD curves (Angstrom^2)."""
    t_ns = np.arange(LAGS) * dt_au * constants.AUTIME2NS
    out = []
    for c in curves:
        slope = least_squares_slope(t_ns[TRIM:-TRIM], c[TRIM:-TRIM])
        out.append(slope * constants.ANG2CM ** 2 * constants.SEC2NS / 6
                   / (constants.KB * temp / constants.EV2J))
    return np.array(out)


def test_fp32_batch_diffusivity_within_two_sigma():
    w = workloads.corundum(4)
    table = w.build_table()
    run = w.make_run(table)
    fam = B.Family(run, table)
    kT = 300.0 * constants.K2AUTEMP
    rate = B.expected_rate_per_carrier(fam, kT)
    dt = HOPS / rate / ROWS
    n_traj = 256
    occs, words = rng.seeded_starts(
        31, n_traj, [fam.hop_sites] * n_traj, [1] * n_traj)
    trajs = [dict(family=0, n_carriers=1, carrier_q=-1.0, field=np.zeros(3),
                  kT=kT, time_interval=dt, n_path=2 * ROWS + 1,
                  max_hops=HOPS) for _ in range(n_traj)]
    bt = B.KmcBatch([fam], "f64", "f32")
    bt.setup(trajs, occs, words, msd_lags=[LAGS] * n_traj,
             msd_scale=1 / constants.ANG2BOHR, waves_per_cu=4)
    bt.run()
    curves = bt.msd()
    res = bt.results()
    assert np.all(res["hops"] == HOPS)
    d_gpu = _d_over_kbt(curves, dt, 300.0)
    mean, sem = d_gpu.mean(), d_gpu.std() / np.sqrt(n_traj)
    _, dcm = exact_diffusivity(run.nbr_site, run.nbr_vec, run.nbr_lambda,
                               run.nbr_vab, run.nbr_count,
                               np.arange(w.n_sites) % 12, run.material.vn,
                               run.temp)
    exact = over_kbt(dcm, 300.0)
    print(f"GPU FP32 D/kBT = {mean:.5e} +- {sem:.2e}; exact {exact:.5e}")
    assert abs(mean - exact) <= 2 * sem

    # FP64 oracle on the first 24 matched seeds
    n_o = 24
    pb = oracle_problem(run, table=table)
    pb.time_interval = dt
    pb.n_path = 2 * ROWS + 1
    d_or = []
    for t in range(n_o):
        gen = random.Random()
        gen.setstate(seeding.trajectory_states(31, n_traj)[t])
        occ = seeding.initial_occupancy(gen, run.hop_site_lists["electron"],
                                        1)
        assert occ == occs[t]
        r = kmc.run(pb, occ, gen.random, max_hops=HOPS)
        P = r["start"]
        pos = r["unwrapped"][:P].reshape(P, 1, 3) / constants.ANG2BOHR
        sd = omsd.squared_displacement(pos, 1, P, LAGS)[0, :, 0]
        d_or.append(_d_over_kbt([sd], dt, 300.0)[0])
    d_or = np.array(d_or)
    m_o, s_o = d_or.mean(), d_or.std() / np.sqrt(n_o)
    m_g = d_gpu[:n_o].mean()
    s_g = d_gpu[:n_o].std() / np.sqrt(n_o)
    print(f"matched seeds: oracle {m_o:.5e} +- {s_o:.2e}, GPU {m_g:.5e} "
          f"+- {s_g:.2e}")
    assert abs(m_o - m_g) <= 2 * np.hypot(s_o, s_g)
    assert abs(m_o - exact) <= 2 * s_o + 1e-300
    bt.close()


def test_fp32_table_batch_diffusivity_within_two_sigma():
    """The same check with the family's table held in FP32 on the device."""
    w = workloads.corundum(4)
    table = w.build_table()
    run = w.make_run(table)
    fam = B.Family(run, table)
    kT = 300.0 * constants.K2AUTEMP
    rate = B.expected_rate_per_carrier(fam, kT)
    dt = HOPS / rate / ROWS
    n_traj = 256
    occs, words = rng.seeded_starts(
        37, n_traj, [fam.hop_sites] * n_traj, [1] * n_traj)
    trajs = [dict(family=0, n_carriers=1, carrier_q=-1.0, field=np.zeros(3),
                  kT=kT, time_interval=dt, n_path=2 * ROWS + 1,
                  max_hops=HOPS) for _ in range(n_traj)]
    bt = B.KmcBatch([fam], "f32", "f32")
    bt.setup(trajs, occs, words, msd_lags=[LAGS] * n_traj,
             msd_scale=1 / constants.ANG2BOHR, waves_per_cu=4)
    bt.run()
    curves = bt.msd()
    assert np.all(bt.results()["hops"] == HOPS)
    bt.close()
    d_gpu = _d_over_kbt(curves, dt, 300.0)
    mean, sem = d_gpu.mean(), d_gpu.std() / np.sqrt(n_traj)
    _, dcm = exact_diffusivity(run.nbr_site, run.nbr_vec, run.nbr_lambda,
                               run.nbr_vab, run.nbr_count,
                               np.arange(w.n_sites) % 12, run.material.vn,
                               run.temp)
    exact = over_kbt(dcm, 300.0)
    print(f"GPU FP32-table D/kBT = {mean:.5e} +- {sem:.2e}; exact {exact:.5e}")
    assert abs(mean - exact) <= 2 * sem


@pytest.mark.parametrize("temp,direction,seed", [
    (250.0, (1, 0, 0), 41),   # along a: fast basal hops, drift ~100 SEM
    (400.0, (0, 0, 1), 43),   # along c: slow c hops, drift ~10 SEM
])
def test_fp32_batch_drift_velocity_within_two_sigma(temp, direction, seed):
    """One electron, |E| = 1e7 V/m along a lattice direction, FP32 rates:
    the mean velocity along the field of 256 trajectories of 20000 hops
    (path rows, start to last full path step) lies within 2 SEM of the
    exact drift velocity of the sublattice walk with the field-shifted
    rates (oracle.diffusion.exact_drift_velocity)."""
    from oracle.diffusion import exact_drift_velocity
    w = workloads.corundum(4)
    table = w.build_table()
    run = w.make_run(table)
    fam = B.Family(run, table)
    kT = temp * constants.K2AUTEMP
    d = np.asarray(direction, dtype=float) @ run.material.lattice_matrix
    d /= np.linalg.norm(d)
    field = 1e7 * constants.VPM2AUFIELD * d
    rate = B.expected_rate_per_carrier(fam, kT, field)
    dt = HOPS / rate / ROWS
    n_traj = 256
    occs, words = rng.seeded_starts(
        seed, n_traj, [fam.hop_sites] * n_traj, [1] * n_traj)
    trajs = [dict(family=0, n_carriers=1, carrier_q=-1.0, field=field,
                  kT=kT, time_interval=dt, n_path=2 * ROWS + 1,
                  max_hops=HOPS) for _ in range(n_traj)]
    bt = B.KmcBatch([fam], "f64", "f32")
    bt.setup(trajs, occs, words, waves_per_cu=4)
    bt.run()
    res = bt.results()
    assert np.all(res["hops"] == HOPS)
    v = []
    for k in range(n_traj):
        rows = bt.rows(k)
        last = int(res["start"][k]) - 1
        v.append((rows[last] - rows[0]) @ d / (last * dt))
    bt.close()
    v = np.array(v)
    mean, sem = v.mean(), v.std() / np.sqrt(n_traj)
    exact = exact_drift_velocity(run.nbr_site, run.nbr_vec, run.nbr_lambda,
                                 run.nbr_vab, run.nbr_count,
                                 np.arange(w.n_sites) % 12, run.material.vn,
                                 kT, field) @ d
    print(f"T={temp} K along {direction}: v = {mean:.5e} +- {sem:.2e} "
          f"bohr/au; exact {exact:.5e}")
    assert abs(mean - exact) <= 2 * sem
    assert abs(exact) > 5 * sem   # the check resolves the drift

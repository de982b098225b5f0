"""The oracle's reductions (MSD species/trajectory means, slopes, D/kBT,
drift mobility, per-hop occupancy) against hand arithmetic, scipy, CODATA
and closed forms, and the product's host reductions against the oracle."""
import random

import numpy as np
import pytest
import scipy.constants as sc
from scipy.stats import linregress

from oracle import kmc, msd as omsd
from oracle.diffusion import exact_diffusivity, over_kbt
from oracle.neighbors import hop_neighbor_list
from pycd_amd import constants, workloads
from pycd_amd.core import Analysis, least_squares_slope
from tests.systems import oracle_problem


def _linear_sd(n_traj_rates, n_lags):
    """sd[t, tau, c] = rate[t][c] * tau."""
    tau = np.arange(n_lags, dtype=float)
    return np.stack([np.stack([r * tau for r in rates], axis=1)
                     for rates in n_traj_rates])


def test_reduce_by_hand_one_species():
    """Two trajectories, two carriers of one species, sd linear in the lag:
    carrier means 3 tau and 2 tau, so every quantity has a closed form."""
    dt, tc, trim, temp = 0.5, 4.0, 1, 300.0
    sd = _linear_sd([[2.0, 4.0], [1.0, 3.0]], 6)
    out = omsd.reduce(sd, [2, 0], ["electron", "hole"], dt, tc, trim, 3, temp)
    tau = np.arange(6)
    assert out["species_types"] == ["electron"]
    assert np.allclose(out["msd_data"][:, 0], tau * 2.0, rtol=0, atol=1e-15)
    assert np.allclose(out["msd_data"][:, 1], 2.5 * tau, rtol=1e-15)
    # std of (3 tau, 2 tau) is tau / 2
    assert np.allclose(out["sem_data"][:, 0], 0.5 * tau / np.sqrt(2),
                       rtol=1e-15)
    # time = 2 tau: slopes 1.5 and 1.0 per time unit
    assert np.allclose(out["slope_data"][:, 0], [1.5, 1.0], rtol=1e-14)
    kbt = constants.KB * temp / constants.EV2J
    d = 1.25 * 1e-16 * 1e9 / 6 / kbt   # A^2/ns -> cm^2/s, / (2 n_dim)
    assert out["diffusivity"][0] == pytest.approx(d, rel=1e-13)
    assert out["diffusivity_sem"][0] == pytest.approx(
        0.25 / np.sqrt(2) * 1e-7 / 6 / kbt, rel=1e-13)


def test_reduce_by_hand_two_species_and_trim():
    """Species 0 = carrier 0, species 1 = carriers 1-2; the trimmed rows
    carry values that would change the slope if they were used."""
    sd = _linear_sd([[1.0, 2.0, 6.0], [3.0, 4.0, 8.0], [5.0, 0.0, 2.0]], 9)
    sd[:, :2] += 100.0
    sd[:, -2:] -= 50.0
    out = omsd.reduce(sd, [1, 2], ["a", "b"], 1.0, 1.0, 2, 3, 300.0)
    assert out["species_types"] == ["a", "b"]
    assert np.allclose(out["slope_data"][:, 0], [1.0, 3.0, 5.0], rtol=1e-13)
    assert np.allclose(out["slope_data"][:, 1], [4.0, 6.0, 1.0], rtol=1e-13)
    tau = np.arange(2, 7)
    assert np.allclose(out["msd_data"][2:7, 2], 11.0 / 3 * tau, rtol=1e-14)


def test_least_squares_slope_matches_linregress():
    rng = np.random.default_rng(5)
    for n in (3, 17, 400):
        x = np.sort(rng.uniform(0, 10, n))
        y = 0.7 * x + rng.normal(size=n)
        ref = linregress(x, y)[0]
        assert least_squares_slope(x, y) == pytest.approx(ref, rel=1e-12)


def test_diffusivity_over_kbt_from_the_exact_slope():
    """The slope of MSD = 2 n_dim D t for the exact corundum D, passed
    through the D/kBT conversion, gives the independently evaluated
    4.394872e-2 cm^2/Vs."""
    w = workloads.corundum(3, hop_lists=hop_neighbor_list)
    run = w.make_run(np.zeros((w.n_sites, w.n_sites)))
    sub = np.arange(w.n_sites) % 12
    _, dcm = exact_diffusivity(run.nbr_site, run.nbr_vec, run.nbr_lambda,
                               run.nbr_vab, run.nbr_count, sub,
                               run.material.vn, run.temp)
    slope = 2 * 3 * dcm / (1e-16 * 1e9)        # cm^2/s -> A^2/ns
    val = omsd.diffusivity_over_kbt(slope, 3, 300.0)
    assert val == pytest.approx(4.394872e-2, rel=2e-6)
    assert val == pytest.approx(over_kbt(dcm, 300.0), rel=1e-14)


def test_analysis_reduction_equals_oracle(tmp_path, hematite):
    """Analysis.reduce_msd (the host half of compute_msd) against the
    oracle on four trajectories of three carriers."""
    rng = np.random.default_rng(11)
    n_traj, M, nc = 4, 60, 3
    sd = np.cumsum(rng.uniform(0, 2, size=(n_traj, M, nc)), axis=1)
    sd[:, 0] = 0
    sim = hematite.sim
    an = Analysis(hematite.material, 3, np.array([nc, 0]), n_traj,
                  sim["t_final"], sim["time_interval"], sim["msd_t_final"],
                  5, sim["temp"], sim["repr_time"], sim["repr_dist"])
    an.num_msd_steps_per_traj = M
    got = an.reduce_msd(sd, tmp_path)
    ref = omsd.reduce(sd, [nc, 0], hematite.material.species_types,
                      an.time_interval, an.time_conversion, 5, 3, sim["temp"])
    assert np.allclose(got.msd_data, ref["msd_data"], rtol=1e-14, atol=0)
    assert np.allclose(got.sem_data, ref["sem_data"], rtol=1e-13, atol=0)
    assert np.allclose(an.slope_data, ref["slope_data"], rtol=1e-12, atol=0)
    assert got.species_types == ref["species_types"]
    log = (tmp_path / f"MSD_Analysis_{got.file_name}.log").read_text()
    for line in omsd.report_lines(ref["species_types"], ref["diffusivity"],
                                  ref["diffusivity_sem"]):
        assert line in log
    saved = np.load(tmp_path / f"MSD_Data_{got.file_name}.npy")
    assert np.array_equal(saved, got.msd_data)


def test_mobility_unit_is_the_atomic_unit_of_mobility():
    """One atomic unit of drift (bohr per au time per au field) is
    e a0^2 / hbar = 4.2544e-2 cm^2/(V s) (CODATA via scipy)."""
    field = np.array([0.0, 0.0, 2.0])
    drift = np.zeros((1, 1, 3))
    drift[0, 0] = [5.0, -1.0, 2.0 * 4.0]          # v.E / |E|^2 = 4 au
    mob, per = kmc.drift_mobility(drift, field, 2.0, [1, 0], ["e", "h"], 1)
    unit = sc.e * sc.physical_constants["Bohr radius"][0] ** 2 / sc.hbar * 1e4
    assert mob[0, 0] == pytest.approx(4 * unit, rel=1e-6)
    assert per[0][0] == "e" and len(per) == 1


def test_mobility_species_means_by_hand():
    field = np.array([1.0, 0.0, 0.0])
    drift = np.zeros((2, 3, 3))
    drift[:, :, 0] = [[1.0, 2.0, 4.0], [3.0, 6.0, 0.0]]
    mob, per = kmc.drift_mobility(drift, field, 1.0, [1, 2], ["a", "b"], 2)
    u = mob[0, 0] / 1.0
    assert np.allclose(mob[:, :], np.array([[1, 2, 4], [3, 6, 0]]) * u)
    # species a: trajectory values 1, 3; species b: carrier means 3, 3
    assert per[0][1] == pytest.approx(2 * u) and per[0][2] == pytest.approx(
        1 * u / np.sqrt(2))
    assert per[1][1] == pytest.approx(3 * u) and per[1][2] == 0.0


def test_run_drift_mobility_equals_oracle(tmp_path):
    """Run.compute_drift_mobility (host) against the oracle, on a run with
    a field along c and two electrons."""
    w = workloads.corundum(3, n_carriers=2, field_vpm=1e6,
                           hop_lists=hop_neighbor_list)
    run = w.make_run(np.zeros((w.n_sites, w.n_sites)))
    rng = np.random.default_rng(3)
    run.n_traj = 3
    drift = rng.normal(size=(3, 2, 3))
    lines = run.compute_drift_mobility(drift, tmp_path, [])
    mob, per = kmc.drift_mobility(drift, run.electric_field,
                                  run.electric_field_mag, run.species_count,
                                  run.material.species_types, 3)
    assert np.allclose(np.loadtxt(tmp_path / "drift_mobility.dat"), mob,
                       rtol=1e-15, atol=0)
    name, mean, sem = per[0]
    assert lines[0] == (f"Estimated value of {name} drift mobility is: "
                        f"{mean:4.3e} cm2/V.s.\n")
    assert lines[1] == (f"Standard error of mean in {name} drift mobility "
                        f"is: {sem:4.3e} cm2/V.s.\n")


def test_occupancy_replay_matches_the_loop(hematite):
    """Occupancy rebuilt from the chosen processes equals the loop's own
    occupancy after every hop (three carriers, 400 hops)."""
    run = hematite.run
    gen = random.Random(7)
    occ = gen.sample(list(run.hop_site_lists["electron"]), 3)
    pb = oracle_problem(run)
    pb.carrier_q = np.full(3, -1.0)
    res = kmc.run(pb, occ, gen.random, max_hops=400, record=True)
    rows = kmc.replay_occupancy(occ, res["procs"], pb.count, pb.nbr)
    assert rows.shape == (401, 3)
    assert np.array_equal(rows, res["occ_rows"])
    moved = np.sum(rows[1:] != rows[:-1], axis=1)
    assert np.all(moved <= 1)
    # the product's rebuild, which a doped run saves as occupancy.npy
    assert np.array_equal(run.occupancy_rows(occ, res["procs"]), rows)

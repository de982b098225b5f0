"""The oracle against the reference's shipped vectors and closed forms."""
import pickle
import random

import numpy as np
import pytest

from oracle import ewald, kmc, seeding, msd
from pycd_amd import constants
from tests.systems import oracle_problem


@pytest.mark.parametrize("name", ["hematite", "bvo"])
def test_table_from_shipped_vectors(name, request):
    ex = request.getfixturevalue(name)
    al, rc, kc = ewald.ewald_units(ex.cfg["alpha"], ex.cfg["r_cut"],
                                   ex.cfg["k_cut"])
    lat = ex.material.lattice_matrix
    size = np.array(ex.sim["system_size"])
    table = ewald.pair_table(ex.pairwise, lat, size, al, rc, kc,
                             ex.cfg["dielectric_constant"])
    # the shipped table was produced by this very arithmetic
    assert np.max(np.abs(table - ex.table)) <= 2e-17 * 8


@pytest.mark.parametrize("name", ["hematite", "bvo"])
def test_table_from_positions(name, request):
    """Fresh minimum images of the shipped site order give the same table."""
    ex = request.getfixturevalue(name)
    lat = ex.material.lattice_matrix
    size = np.array(ex.sim["system_size"])
    cell = ewald.simulation_cell(lat, size)
    pw = ewald.pairwise_vectors(ex.shipped_pos, cell, ex.sim["pbc"])
    r_new = np.linalg.norm(pw, axis=2)
    r_old = np.linalg.norm(ex.pairwise, axis=2)
    assert np.max(np.abs(r_new - r_old)) < 1e-13
    al, rc, kc = ewald.ewald_units(ex.cfg["alpha"], ex.cfg["r_cut"],
                                   ex.cfg["k_cut"])
    table = ewald.pair_table(pw, lat, size, al, rc, kc,
                             ex.cfg["dielectric_constant"])
    assert np.max(np.abs(table - ex.table)) < 1e-15 * np.max(np.abs(ex.table))


@pytest.mark.parametrize("name", ["hematite", "bvo"])
def test_table_rows_from_positions(name, request):
    """`table_rows` (the large-N form: its own minimum images and a matrix
    product over the k half-space) reproduces every row of the shipped
    table from the shipped site order."""
    ex = request.getfixturevalue(name)
    al, rc, kc = ewald.ewald_units(ex.cfg["alpha"], ex.cfg["r_cut"],
                                   ex.cfg["k_cut"])
    n = len(ex.shipped_pos)
    rows = ewald.table_rows(ex.shipped_pos, range(n),
                            ex.material.lattice_matrix,
                            np.array(ex.sim["system_size"]), ex.sim["pbc"],
                            al, rc, kc, ex.cfg["dielectric_constant"])
    assert np.max(np.abs(rows - ex.table)) <= 1e-15


def test_half_space_is_a_half():
    km = np.array([2, 1, 3])
    kv = ewald.half_space_kvectors(km)
    total = int(np.prod(2 * km + 1)) - 1
    assert len(kv) == total // 2
    keys = {tuple(v) for v in kv}
    for v in kv:
        assert tuple(-v) not in keys
        nz = v[np.nonzero(v)[0][0]]
        assert nz < 0  # first non-zero component negative (mgrid order)


def test_minimum_image_vs_27_images():
    """Same check as the reference's tests/test_pbc_optimization.py:43-85."""
    rng = np.random.default_rng(0)
    lat = np.array([[3.0, 0.1, 0.0], [0.1, 3.0, 0.0], [0.0, 0.0, 3.0]])
    size = np.array([2, 2, 2])
    cell = ewald.simulation_cell(lat, size)
    d = rng.uniform(-8, 8, size=(500, 3))
    got = ewald.mic_rows(d, cell, [1, 1, 1])
    shifts = np.array([[x, y, z] for x in (-1, 0, 1) for y in (-1, 0, 1)
                       for z in (-1, 0, 1)]) @ cell
    frac = d @ np.linalg.inv(cell)
    wrapped = (frac - np.round(frac)) @ cell
    cand = wrapped[:, None, :] + shifts[None, :, :]
    best = np.min(np.linalg.norm(cand, axis=2), axis=1)
    assert np.allclose(np.linalg.norm(got, axis=1), best, atol=1e-10)


def test_ewald_total_energy_identity():
    """Charge-neutral lattice: q^T A q equals the textbook Ewald energy.

    For point charges q_i the Ewald energy with splitting parameter eta
    (erfc(eta r)) is E = 1/2 sum_{i!=j} q_i q_j erfc(eta r)/r
    + (2 pi / V) sum_{k != 0} exp(-k^2/(4 eta^2))/k^2 |S(k)|^2
    - eta/sqrt(pi) sum q_i^2.  The table holds half pair potentials
    (E = q^T A q), uses eta^2 = alpha and puts 0.5 on the real-space
    diagonal, so q^T A q = E + 0.5 sum q_i^2 (real-space minimum image
    only, r_cut below half the box)."""
    rng = np.random.default_rng(1)
    lat = np.diag([7.0, 8.0, 9.0])
    size = np.array([1, 1, 1])
    n = 6
    pos = rng.uniform(0, 1, size=(n, 3)) @ lat
    q = np.array([1.0, -1.0, 2.0, -2.0, 0.5, -0.5])
    alpha, r_cut, k_cut = 0.45, 3.4, 4.0
    cell = ewald.simulation_cell(lat, size)
    pw = ewald.pairwise_vectors(pos, cell, [1, 1, 1])
    table = ewald.pair_table(pw, lat, size, alpha, r_cut, k_cut, 1.0)
    eta = np.sqrt(alpha)
    r = np.linalg.norm(pw, axis=2)
    from scipy.special import erfc
    e_real = 0.0
    for i in range(n):
        for j in range(n):
            if i != j and r[i, j] < r_cut:
                e_real += 0.5 * q[i] * q[j] * erfc(eta * r[i, j]) / r[i, j]
    vol = np.prod(np.diag(lat))
    e_rec = 0.0
    b = 2 * np.pi / np.diag(lat)
    km = np.ceil(k_cut / b).astype(int)
    for a in range(-km[0], km[0] + 1):
        for bb in range(-km[1], km[1] + 1):
            for c in range(-km[2], km[2] + 1):
                if a == bb == c == 0:
                    continue
                k = np.array([a, bb, c]) * b
                k2 = k @ k
                if k2 >= k_cut ** 2:
                    continue
                s = np.sum(q * np.exp(1j * (pos @ k)))
                e_rec += (2 * np.pi / vol) * np.exp(-k2 / (4 * alpha)) / k2 \
                    * abs(s) ** 2
    e_self = -eta / np.sqrt(np.pi) * np.sum(q * q)
    lhs = q @ table @ q
    rhs = (e_real + e_rec + e_self) + 0.5 * np.sum(q * q)
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(rhs))


@pytest.mark.parametrize("name", ["hematite", "bvo"])
def test_seeding_protocol_bytes(name, request):
    """preproduction's state equals the shipped dump byte for byte."""
    ex = request.getfixturevalue(name)
    state = seeding.trajectory_states(ex.sim["random_seed"], 1)[0]
    assert pickle.dumps(state, protocol=3) == ex.rnd_dump


def _replay(ex, max_hops=None):
    pb = oracle_problem(ex.run, table=ex.table)
    state = seeding.trajectory_states(ex.sim["random_seed"], 1)[0]
    gen = random.Random()
    gen.setstate(state)
    occ = seeding.initial_occupancy(gen, ex.run.hop_site_lists["electron"], 1)
    return kmc.run(pb, occ, gen.random, max_hops=max_hops, record=True)


def test_hematite_replay_all_hops(hematite):
    """Every shipped hop time and the whole unwrapped path (F5)."""
    res = _replay(hematite)
    td = hematite.time_data
    assert res["hops"] == len(td) - 1
    rel = np.abs(res["time_data"][1:] - td[1:]) / td[1:]
    assert rel.max() < 1e-12
    assert np.max(np.abs(res["unwrapped"] - hematite.unwrapped)) < 1e-9


def test_bvo_replay_all_hops(bvo):
    """All hop times; path equal up to whole supercell translations (the
    2x2x1 cell has hop vectors exactly at half a box edge, where the
    minimum-image choice is a floating-point tie)."""
    res = _replay(bvo)
    td = bvo.time_data
    assert res["hops"] == len(td) - 1
    rel = np.abs(res["time_data"][1:] - td[1:]) / td[1:]
    assert rel.max() < 1e-12
    cell = bvo.neighbors.simulation_cell_matrix
    frac = (res["unwrapped"] - bvo.unwrapped).reshape(-1, 3) @ np.linalg.inv(
        cell)
    assert np.max(np.abs(frac - np.round(frac))) < 1e-9


def test_delta_g0_is_energy_difference(hematite):
    """dG0 of every process equals E(after) - E(before), E = q^T A q: the
    identity the reference's term01/term02 expression implements."""
    ex = hematite
    run = ex.run
    A = ex.table
    q = run.base_charge_config()[:, 0].copy()
    occ = [int(run.hop_site_lists["electron"][7])]
    q[occ[0]] += -1.0
    pb = oracle_problem(run, table=A)
    gen = iter([0.3, 0.6])
    res = kmc.run(pb, occ, lambda: next(gen), max_hops=1, record=True)
    e0 = q @ A @ q
    s = occ[0]
    for j in range(run.nbr_count[s]):
        n = run.nbr_site[s, j]
        q2 = q.copy()
        q2[s] += 1.0
        q2[n] -= 1.0
        term01 = np.dot(q, A[n] - A[s])
        term02 = -1.0 * (A[s, s] - A[s, n])
        dg0 = 2 * -1.0 * (term01 + term02)
        assert abs(dg0 - (q2 @ A @ q2 - e0)) < 1e-12
    assert res["hops"] == 1


def test_detailed_balance(hematite):
    """Zero field, equal lambda: k(i->j)/k(j->i) = exp(-dG0/kT)."""
    ex = hematite
    run = ex.run
    A = ex.table
    q = run.base_charge_config()[:, 0]
    kT = run.temp
    s = int(run.hop_site_lists["electron"][3])
    for j in range(run.nbr_count[s]):
        n = int(run.nbr_site[s, j])
        lam, vab = run.nbr_lambda[s, j], run.nbr_vab[s, j]

        def rate(a, b):
            qq = q.copy()
            qq[a] += -1.0
            dg = 2 * -1.0 * (np.dot(qq, A[b] - A[a]) - (A[a, a] - A[a, b]))
            return np.exp(-(((lam + dg) ** 2 / (4 * lam)) - vab) / kT), dg

        k_f, dg = rate(s, n)
        k_b, _ = rate(n, s)
        assert np.isclose(k_f / k_b, np.exp(-dg / kT), rtol=1e-9)


def test_msd_oracle_brute_force():
    rng = np.random.default_rng(2)
    P, nc, M = 40, 3, 12
    pos = np.cumsum(rng.normal(size=(P, nc, 3)), axis=0)
    sd = msd.squared_displacement(pos, 1, P, M)
    for c in range(nc):
        assert np.allclose(sd[0, :, c], msd.brute_force_sd(pos[:, c], M),
                           rtol=1e-13, atol=0)


def test_msd_oracle_ballistic():
    P, M = 50, 20
    v = np.array([0.3, -0.2, 0.7])
    pos = (np.arange(P)[:, None] * v)[:, None, :]
    sd = msd.squared_displacement(pos, 1, P, M)
    tau = np.arange(M)
    assert np.allclose(sd[0, :, 0], (v @ v) * tau ** 2, rtol=1e-13)


def test_half_space_order_matches_product():
    from pycd_amd.core import System
    for km in ([2, 1, 3], [0, 0, 2], [1, 2, 0], [3, 3, 3]):
        assert np.array_equal(System.half_space(np.array(km)),
                              ewald.half_space_kvectors(np.array(km)))


@pytest.mark.parametrize("name", ["hematite", "bvo"])
def test_k_vector_energies(name, request):
    """Per-k energies of the lattice charges: the product's structure-factor
    form |S(k)|^2 equals the reference's double sum over the shipped
    minimum-image vectors, and twice their sum is q^T A_k q (the
    reciprocal part of the table)."""
    ex = request.getfixturevalue(name)
    al, rc, kc = ewald.ewald_units(ex.cfg["alpha"], ex.cfg["r_cut"],
                                   ex.cfg["k_cut"])
    lat = ex.material.lattice_matrix
    size = np.array(ex.sim["system_size"])
    _, volume, recip = ewald.geometry(lat, size)
    km = ewald.k_max_for(kc, recip)
    system = ex.system
    q = system.base_charge_config_for_accuracy_analysis("full")[:, 0]
    kv_o, e_o = ewald.k_vector_energies(ex.pairwise, recip, volume, km, al,
                                        kc, q)
    # the product's own geometry, from the site order it built
    kv_p, e_p = system.k_vector_energies(q, al, kc, km)
    assert np.array_equal(kv_p, kv_o)
    assert np.allclose(e_p, e_o, rtol=1e-9, atol=1e-12 * np.max(np.abs(e_o)))
    full = ewald.pot_k(ex.pairwise, recip, volume, km, al, kc)
    assert 2 * e_o.sum() == pytest.approx(q @ full @ q, rel=1e-12)

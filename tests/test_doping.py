"""Dopant placement and its per-site effects (host logic, no GPU).

No reference fixture holds a doped run (the shipped examples have
num_dopants [0, 0]), so these checks pin the placement by what fixes it
independently:
  * shells are hop distances: scipy's breadth-first shortest paths over the
    hop graph built from the per-site slot table;
  * the shell order within a shell is the reference's own walk (every inner
    site, list membership), restated literally on the small lattice;
  * random insertion keeps dopants more than min_shell_separation hops
    apart, and the same seed places the same dopants;
  * the shell table gives each site its distance to the nearest dopant, the
    earlier dopant on ties; energies, charges and the initial occupancy
    follow from it.
"""
import random

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

from pycd_amd import constants, doping
from tests.systems import DOPING_SHIFTS as SHIFTS, doped_run as _doped_run


def _hop_distances(run):
    """All-pairs hop distance over the Fe sites (scipy BFS)."""
    n = run.nbr_site.shape[0]
    rows, cols = [], []
    for s in range(n):
        for j in range(run.nbr_count[s]):
            rows.append(s)
            cols.append(int(run.nbr_site[s, j]))
    g = csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    return shortest_path(g, unweighted=True)


def _reference_walk(center, n_shells, targets):
    """The reference's shell walk as written: every inner site each shell,
    list membership (PyCD/core.py:2352-2389)."""
    out, inner = [], []
    for k in range(n_shells + 1):
        if k == 0:
            cur = [center]
        else:
            walk = []
            for s in inner:
                walk.extend(targets(s))
            cur = [x for x in walk if x not in inner]
        cur = list(set(cur))
        inner.extend(cur)
        out.append(cur)
    return out


@pytest.fixture(scope="module")
def fe(hematite):
    run = _doped_run(hematite, 3)
    fe_sites = [int(s) for s in run.hop_site_lists["electron"]]
    return run, fe_sites, _hop_distances(run)


def test_shells_are_hop_distances(fe):
    run, fe_sites, dist = fe
    targets = run.hop_targets(0)
    for c in fe_sites[::7]:
        ecc = int(np.max(dist[c, fe_sites]))
        sh = run.get_shell_based_neighbors(c, ecc + 1)
        for k, members in enumerate(sh):
            want = {s for s in fe_sites if dist[c, s] == k}
            assert set(members) == want and len(members) == len(want)
        assert sh[-1] == []
        assert sh == _reference_walk(c, ecc + 1, targets)
        assert len(doping.all_shells(c, targets)) == ecc + 1


def test_shell_lookup_bins(fe):
    run, fe_sites, dist = fe
    look = run.shell_lookup("Fe")
    sample = fe_sites[0]
    assert look.max_shells == int(np.max(dist[sample, fe_sites]))
    assert np.all(np.diff(look.dist_bins) >= 0)
    # a distance held by sites of one shell only maps to that shell
    by_dist = {}
    for s in fe_sites:
        d = round(run.neighbors.compute_distance(run.system_size, sample, s)
                  / constants.ANG2BOHR, 9)
        by_dist.setdefault(d, set()).add(int(dist[sample, s]))
    checked = 0
    for d, ks in by_dist.items():
        if len(ks) == 1:
            assert look.shell(d) == next(iter(ks))
            checked += 1
    assert checked >= 3


def test_random_insertion_separation_and_seed(fe, hematite):
    run, fe_sites, dist = fe
    random.seed(11)
    a = run.get_doping_distribution()
    random.seed(11)
    b = run.get_doping_distribution()
    assert a == b and len(a["Ti"]) == 3
    for i, s in enumerate(a["Ti"]):
        for t in a["Ti"][i + 1:]:
            assert dist[s, t] > 1
    # more dopants than the separation leaves room for: fewer placed
    crowded = _doped_run(hematite, 40, sep=3)
    random.seed(3)
    got = crowded.get_doping_distribution()["Ti"]
    assert 1 <= len(got) < 40
    for i, s in enumerate(got):
        for t in got[i + 1:]:
            assert dist[s, t] > 3


def test_preproduction_shell_table_and_effects(fe, hematite, tmp_path):
    run, fe_sites, dist = fe
    run.preproduction(tmp_path, 5)
    tables = []
    for k in range(run.n_traj):
        d = tmp_path / f"traj{k + 1}"
        table = np.load(d / "site_indices.npy")
        tables.append(table)
        log = (d / "PreProduction.log").read_text()
        assert log.startswith(f"Trajectory {k + 1}:\nInserted 3 sites of "
                              f"dopant element type Ti\n")
        assert table.shape == (len(fe_sites), 4)
        assert table[:, 0].tolist() == fe_sites
        dopants = [int(s) for s in table[table[:, 3] == 0, 0]]
        assert len(dopants) == 3
        # each site: hop distance to the nearest dopant, earliest on ties
        order = [int(table[table[:, 0] == s, 2][0]) for s in dopants]
        inserted = [s for _, s in sorted(zip(order, dopants))]
        for row in table:
            dd = [dist[s, row[0]] for s in inserted]
            assert row[3] == min(dd)
            assert row[2] == int(np.argmin(dd))
            assert row[1] == 0
        # energies, charges, occupancy of this trajectory
        energy, q, sites = run.doped_inputs(d)
        shift = np.zeros(run.nbr_site.shape[0])
        for row in table:
            if row[3] < len(SHIFTS):
                shift[row[0]] = SHIFTS[row[3]] * constants.EV2HARTREE
        assert np.array_equal(energy,
                              run.undoped_system_relative_energies + shift)
        base = run.base_charge_config()[:, 0]
        assert np.all(q[dopants] == 4.0)
        rest = np.setdiff1d(np.arange(len(q)), dopants)
        assert np.array_equal(q[rest], base[rest])
        assert sites == {"Ti": sorted(dopants)}
        random.seed(k)
        occ = run.generate_initial_occupancy(sites)
        assert len(occ) == 1 and occ[0] in dopants
    # trajectories get distinct placements
    assert len({tuple(t[t[:, 3] == 0, 0]) for t in tables}) == run.n_traj


def test_manual_insertion_and_no_initiation(hematite, tmp_path):
    fe_sites = [int(s) for s in
                hematite.run.hop_site_lists["electron"]]
    manual = [fe_sites[4], fe_sites[30]]
    run = _doped_run(hematite, 2, kind="manual", initiation="no",
                     manual=manual)
    assert run.get_doping_distribution() == {"Ti": manual}
    run.preproduction(tmp_path, 2)
    table = np.load(tmp_path / "traj1" / "site_indices.npy")
    assert sorted(table[table[:, 3] == 0, 0].tolist()) == sorted(manual)
    _, _, sites = run.doped_inputs(tmp_path / "traj1")
    random.seed(0)
    occ = run.generate_initial_occupancy(sites)
    random.seed(0)
    assert occ == hematite.run.generate_initial_occupancy({})


def test_unsupported_insertion_types_raise(hematite):
    for kind in ("pairwise", "gradient"):
        with pytest.raises(NotImplementedError):
            _doped_run(hematite, 2, kind=kind)


def test_undoped_run_is_unchanged(hematite):
    run = hematite.run
    assert run.doping_active == 0
    assert np.array_equal(run.system_relative_energies,
                          run.undoped_system_relative_energies)

"""HIP plain KMC path vs the shipped trajectories and the oracle."""
import numpy as np
import pytest

from pycd_amd import _native, workloads
from tests.kmc_helpers import run_gpu, run_oracle, native_problem

pytestmark = pytest.mark.gpu

TIME_TOL = 1e-12  # relative, per hop


def test_hematite_every_shipped_hop(hematite):
    res, _ = run_gpu(hematite.run, hematite.table, 2, 1, record=False)
    td = hematite.time_data
    assert res["hops"] == len(td) - 1
    assert len(res["time_data"]) == len(td)
    rel = np.abs(res["time_data"][1:] - td[1:]) / td[1:]
    print("hematite hops", res["hops"], "max rel time dev", rel.max())
    assert rel.max() < TIME_TOL
    assert np.max(np.abs(res["unwrapped"] - hematite.unwrapped)) < 1e-9


def test_bvo_every_shipped_hop(bvo):
    res, _ = run_gpu(bvo.run, bvo.table, 2, 1, record=False, chunk=1000)
    td = bvo.time_data
    assert res["hops"] == len(td) - 1
    rel = np.abs(res["time_data"][1:] - td[1:]) / td[1:]
    assert rel.max() < TIME_TOL
    cell = bvo.neighbors.simulation_cell_matrix
    frac = (res["unwrapped"] - bvo.unwrapped).reshape(-1, 3) @ np.linalg.inv(
        cell)
    assert np.max(np.abs(frac - np.round(frac))) < 1e-9


def _compare(gpu, ora, n):
    assert gpu["hops"] == ora["hops"] == n
    assert np.array_equal(gpu["procs"], ora["procs"])
    rel = np.abs(gpu["time_data"][1:] - ora["time_data"][1:]) / \
        ora["time_data"][1:]
    assert rel.max() < TIME_TOL
    assert np.max(np.abs(gpu["unwrapped"] - ora["unwrapped"])) < 1e-8
    assert gpu["occupancy"] == ora["occupancy"]
    scale = max(1.0, abs(ora["energy_now"]))
    assert abs(gpu["energy_now"] - ora["energy_now"]) < 1e-10 * scale
    return rel.max()


def test_small_tables_keep_full_rows():
    """Below 32 MB the whole table stays in the caches: no cell-0 rows."""
    w = workloads.corundum(3, n_carriers=4)
    table = w.build_table()
    sess = _native.KmcSession(native_problem(w.make_run(table), table, True))
    assert sess.table_rows() == w.n_sites
    sess.close()


def test_table_that_is_not_periodic_keeps_full_rows():
    """One entry off its cell-0 translate by 1e-3 Ha: the check at creation
    rejects the cell-0 rows and the loop reads the full table."""
    w = workloads.corundum(6, n_carriers=4)
    table = w.build_table().copy()
    run = w.make_run(table)
    table[100, 2000] += 1e-3
    sess = _native.KmcSession(native_problem(run, table, True))
    assert sess.table_rows() == w.n_sites
    sess.close()
    table[100, 2000] -= 1e-3
    sess = _native.KmcSession(native_problem(run, table, True))
    assert sess.table_rows() == 12
    sess.close()


@pytest.mark.parametrize("edge", [6, 8])
def test_cell0_rows_give_the_full_table_events(edge):
    """A periodic supercell's hop loop gathers from the table rows of cell 0
    (the table is translation-invariant): the same events as the loop over
    the full table, the clock within 1e-12 relative."""
    w = workloads.corundum(edge, n_carriers=16, field_vpm=1e6)
    table = w.build_table()
    run = w.make_run(table)
    for compact, rows in ((True, 12), (False, w.n_sites)):
        sess = _native.KmcSession(native_problem(run, table, compact))
        assert sess.table_rows() == rows
        sess.close()
    n = 20000
    comp, _ = run_gpu(run, table, 2, 16, max_hops=n, compact=True)
    full, _ = run_gpu(run, table, 2, 16, max_hops=n, compact=False)
    print(f"edge {edge} max rel time dev", _compare(comp, full, n))


def test_small_corundum_1e5_hops_vs_oracle():
    """3x3x3 (N = 324), 1 electron, zero field: 100000 hops, same events."""
    w = workloads.corundum(3, n_carriers=1)
    table = w.build_table()
    run = w.make_run(table)
    n = 100_000
    gpu, _ = run_gpu(run, table, 2, 1, max_hops=n)
    ora = run_oracle(run, table, 2, 1, max_hops=n)
    print("cfg1 max rel time dev", _compare(gpu, ora, n))


def test_large_corundum_field_prefix_vs_oracle():
    """8x8x8 (N = 6144), 16 electrons, 1e6 V/m along c: first 3000 hops."""
    w = workloads.corundum(8, n_carriers=16, field_vpm=1e6, t_final=2e-8)
    table = w.build_table()
    run = w.make_run(table)
    n = 3000
    gpu, _ = run_gpu(run, table, 2, 16, max_hops=n)
    ora = run_oracle(run, table, 2, 16, max_hops=n)
    print("cfg2 prefix max rel time dev", _compare(gpu, ora, n))
    assert np.allclose(gpu["drift"], ora["drift"], rtol=1e-9, atol=0)


def test_large_corundum_1e6_hops_invariants():
    """Full 1e6-hop run: energy bookkeeping equals q^T A q of the final
    state, clock strictly increasing, path rows consistent."""
    w = workloads.corundum(8, n_carriers=16, field_vpm=1e6, t_final=2e-8)
    table = w.build_table()
    run = w.make_run(table)
    n = 1_000_000
    gpu, occ0 = run_gpu(run, table, 2, 16, max_hops=n, record=False)
    assert gpu["hops"] == n
    t = gpu["time_data"]
    assert np.all(np.diff(t) > 0)
    q = run.base_charge_config()[:, 0].copy()
    np.add.at(q, np.asarray(gpu["occupancy"]), -1.0)
    e_final = q @ table @ q
    q0 = run.base_charge_config()[:, 0].copy()
    np.add.at(q0, np.asarray(occ0), -1.0)
    e0 = q0 @ table @ q0
    drift = abs((gpu["energy_now"] - gpu["energy"][0]) - (e_final - e0))
    print("energy bookkeeping drift over 1e6 hops", drift)
    assert drift < 1e-8
    assert gpu["start"] == int(t[-1] / run.time_interval)

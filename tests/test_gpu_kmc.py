"""HIP plain KMC path vs the shipped trajectories and the oracle."""
import numpy as np
import pytest

from pycd_amd import _native, workloads
from tests.kmc_helpers import run_gpu, run_oracle, native_problem

pytestmark = pytest.mark.gpu

TIME_TOL = 1e-12  # relative, per hop
# dG0 is a difference of potentials of order 1 Ha: the GPU forms q.A[x, :]
# as phi_L + psi instead of the reference's N-long dot, which rounds
# differently at the 1e-15 Ha level whatever the size of dG0 itself
DG0_TOL = 1e-13  # Ha, absolute


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
    assert gpu["start"] == ora["start"]
    # energy_traj.npy and delG0_traj.npy contents: the energy after the
    # last hop of each path step, and the dG0 of the hop that ended it
    escale = max(1.0, np.max(np.abs(ora["energy"])))
    assert np.max(np.abs(gpu["energy"] - ora["energy"])) < 1e-10 * escale
    assert np.max(np.abs(gpu["delg_0"] - ora["delg_0"])) < DG0_TOL
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
    # a NaN entry fails the check too (it is not dropped by the maximum)
    keep = table[100, 2000]
    table[100, 2000] = np.nan
    sess = _native.KmcSession(native_problem(run, table, True))
    assert sess.table_rows() == w.n_sites
    sess.close()
    table[100, 2000] = keep


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


def test_large_corundum_1e6_hops_vs_oracle_record():
    """8x8x8 (N = 6144), 16 electrons, 1e6 V/m along c, 1e6 hops: the
    chosen process of every hop equals the oracle's (BLAKE2b-64 digest of
    each 1000-hop block, tests/golden/cfg2_events.npz, written by
    tests/golden/make_cfg2_events.py from the oracle's own table and
    trajectory); the clock after every 1000th hop within 1e-12 relative;
    every 100th path row, energy and dG0 entry and the final state as the
    oracle's; and the energy bookkeeping equal to q^T A q of the final
    state."""
    from tests.conftest import GOLDEN
    from tests.golden.make_cfg2_events import block_hashes
    fx = np.load(GOLDEN / "cfg2_events.npz")
    w = workloads.corundum(8, n_carriers=16, field_vpm=1e6, t_final=2e-8)
    table = w.build_table()
    run = w.make_run(table)
    n = int(fx["hops"])
    assert n == 1_000_000
    gpu, occ0 = run_gpu(run, table, 2, 16, max_hops=n, record=True)
    assert list(occ0) == fx["initial_occupancy"].tolist()
    assert gpu["hops"] == n
    bad = np.nonzero(block_hashes(gpu["procs"]) != fx["block_hash"])[0]
    assert len(bad) == 0, f"event blocks differ from the oracle's: {bad[:5]}"
    tk = gpu["time_data"][::int(fx["block"])]
    rel = np.abs(tk[1:] - fx["time_k"][1:]) / fx["time_k"][1:]
    print("cfg2 1e6 hops: clock every 1000 hops, max rel dev", rel.max())
    assert rel.max() < TIME_TOL
    assert abs(gpu["time_data"][-1] - fx["t"]) <= TIME_TOL * fx["t"]
    idx = fx["row_index"]
    assert np.max(np.abs(gpu["unwrapped"][idx] - fx["rows_k"])) < 1e-8
    escale = max(1.0, np.max(np.abs(fx["energy_k"])))
    assert np.max(np.abs(gpu["energy"][idx] - fx["energy_k"])) < 1e-9 * escale
    assert np.max(np.abs(gpu["delg_0"][idx] - fx["delg_k"])) < DG0_TOL
    assert gpu["occupancy"] == fx["occupancy"].tolist()
    assert gpu["start"] == int(fx["start"])
    assert abs(gpu["energy_now"] - float(fx["energy_now"])) < 1e-9 * max(
        1.0, abs(float(fx["energy_now"])))
    assert np.allclose(gpu["drift"], fx["drift"], rtol=1e-9, atol=0)
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


def _run_variant(run, table, nc, variant, n, compact=True):
    sess = _native.KmcSession(native_problem(run, table, compact))
    sess.set_variant(variant)
    assert sess.variant() == variant
    res, _ = run_gpu(run, table, 2, nc, max_hops=n, session=sess)
    sess.close()
    return res


@pytest.mark.parametrize("edge,nc,field,compact", [
    (8, 16, 1e6, True),     # 1 process per lane, cell-0 rows
    (4, 32, 1e7, False),    # 2 per lane, full table
    (6, 48, 0.0, True),     # 3 per lane, cell-0 rows
])
def test_wave_kernel_equals_workgroup_kernel(edge, nc, field, compact):
    """The one-wavefront kernel (table entries of every carrier cached in
    LDS, one gather per row and hop) and the workgroup kernel (every entry
    gathered each hop) give the same trajectory bit for bit."""
    w = workloads.corundum(edge, n_carriers=nc, field_vpm=field,
                           t_final=2e-8)
    table = w.build_table()
    run = w.make_run(table)
    n = 4000
    a = _run_variant(run, table, nc, "wave", n, compact)
    b = _run_variant(run, table, nc, "workgroup", n, compact)
    assert a["hops"] == b["hops"] == n
    assert np.array_equal(a["procs"], b["procs"])
    for key in ("time_data", "unwrapped", "energy", "delg_0", "drift"):
        assert np.array_equal(a[key], b[key]), key
    assert a["occupancy"] == b["occupancy"]
    assert a["energy_now"] == b["energy_now"] and a["start"] == b["start"]


def test_wave_kernel_applies_where_it_fits():
    """n_c z <= 256 and the cache within LDS: the wavefront kernel by
    default; 64 electrons with 4 slots each (a 164 KB cache) keep the
    workgroup kernel, and asking for the wavefront kernel is refused."""
    w = workloads.corundum(4, n_carriers=16)
    table = w.build_table()
    run = w.make_run(table)
    sess = _native.KmcSession(native_problem(run, table))
    assert sess.variant() == "wave"
    sess.close()
    w = workloads.corundum(4, n_carriers=64)
    table = w.build_table()
    run = w.make_run(table)
    sess = _native.KmcSession(native_problem(run, table))
    assert sess.variant() == "workgroup"
    with pytest.raises(_native.NativeError):
        sess.set_variant("wave")
    sess.close()

"""Serving loop: continuous batching, table cache, admission by footprint."""
import numpy as np
import pytest

from pycd_amd import workloads
from pycd_amd.serve import KmcServer, Job, RowArena, poisson_jobs

pytestmark = pytest.mark.gpu

_FAMILIES = {}


def corundum_family(edge):
    if edge not in _FAMILIES:
        w = workloads.corundum(edge)
        table = w.build_table()
        _FAMILIES[edge] = (w.make_run(table), table)
    return _FAMILIES[edge]


def _result(job):
    r = job.result
    return (r["t"], r["energy"], r["hops"], r["start"], r["rows"].tobytes(),
            r["msd"].tobytes())


def test_same_alone_and_in_the_mix():
    probe = dict(edge=3, n_carriers=6, hops=3000, seed=77, temp=320.0,
                 field_vpm=2e6)
    alone = KmcServer(corundum_family, n_slots=8, nc_max=16,
                      rows_capacity=1 << 20, slice_hops=3000,
                      rate_precision="f32", path_rows=50, msd_lags=20)
    a = Job(0, **probe)
    alone.submit(a)
    alone.drain()
    alone.close()
    mixed = KmcServer(corundum_family, n_slots=64, nc_max=16,
                      rows_capacity=1 << 22, slice_hops=257,
                      rate_precision="f32", path_rows=50, msd_lags=20)
    rng = np.random.default_rng(0)
    jobs = [Job(k + 1, edge=int(rng.integers(3, 5)),
                n_carriers=int(rng.integers(1, 17)),
                hops=int(rng.integers(500, 4000)), seed=100 + k)
            for k in range(40)]
    b = Job(99, **probe)
    for j in jobs[:20]:
        mixed.submit(j)
    mixed.submit(b)
    for j in jobs[20:]:
        mixed.submit(j)
    mixed.drain()
    mixed.close()
    # a job ends at its hop budget or when its path rows are full
    assert 0 < a.result["hops"] <= 3000
    assert _result(a) == _result(b)
    assert all(j.result is not None for j in jobs)


def test_table_budget_forces_eviction():
    """Compact families of 324 and 768 sites (160.4 kB and 375.3 kB on the
    device with their hop differences) do not fit a 400 kB budget together:
    sizes alternate, so families are evicted and rebuilt, and every job
    still completes.  A job whose family alone exceeds the budget (1500
    sites, 729.6 kB) is rejected instead of waiting forever."""
    srv = KmcServer(corundum_family, n_slots=16, nc_max=8,
                    rows_capacity=1 << 20, table_budget_bytes=400_000,
                    slice_hops=500, path_rows=20, msd_lags=5)
    jobs = [Job(k, edge=3 if k % 2 == 0 else 4, n_carriers=2, hops=1000,
                seed=k) for k in range(8)]
    big = Job(100, edge=5, n_carriers=2, hops=100, seed=5)
    for j in jobs:
        srv.submit(j)
    srv.submit(big)
    assert srv.table_bytes(3) + srv.table_bytes(4) > 400_000
    assert max(srv.table_bytes(3), srv.table_bytes(4)) <= 400_000
    assert srv.table_bytes(5) > 400_000
    srv.drain()
    assert all(j.result is not None and 0 < j.result["hops"] <= 1000
               for j in jobs)
    assert "error" in big.result
    assert srv.evictions >= 1
    srv.close()


def test_row_arena():
    a = RowArena(100)
    x = a.alloc(40)
    y = a.alloc(40)
    assert (x, y) == (0, 40)
    assert a.alloc(30) is None
    a.release(x, 40)
    assert a.alloc(30) == 0
    a.release(y, 40)
    a.release(0, 30)
    assert a.used == 0 and a.free == [(0, 100)]


def test_realtime_poisson_small():
    jobs = poisson_jobs(30, rate=200.0, edges=(3, 4), nc=(1, 8),
                        hops=(1e3, 1e4), seed=3)
    srv = KmcServer(corundum_family, n_slots=64, nc_max=8,
                    rows_capacity=1 << 22, slice_hops=1000, path_rows=50,
                    msd_lags=10)
    stats = srv.serve(jobs)
    srv.close()
    print(stats)
    assert stats["jobs"] == 30
    assert stats["p99_latency_ms"] >= stats["p50_latency_ms"] > 0


def test_cancel_and_timeout_requeue_rerun_identically():
    """Failures forced from the host between launches: a cancelled job
    frees its slot; a re-queued job and the jobs of a slice the watchdog
    declares timed out start over and end with the results of an
    undisturbed run."""
    kw = dict(n_slots=8, nc_max=8, rows_capacity=1 << 20, slice_hops=400,
              rate_precision="f32", path_rows=30, msd_lags=10)
    specs = [dict(edge=3, n_carriers=3, hops=2000, seed=11),
             dict(edge=4, n_carriers=5, hops=1500, seed=12),
             dict(edge=3, n_carriers=2, hops=2500, seed=13)]
    clean = KmcServer(corundum_family, **kw)
    ref = [Job(k, **sp) for k, sp in enumerate(specs)]
    for j in ref:
        clean.submit(j)
    clean.drain()
    clean.close()

    srv = KmcServer(corundum_family, **kw)
    jobs = [Job(k, **sp) for k, sp in enumerate(specs)]
    extra = Job(9, edge=3, n_carriers=4, hops=5000, seed=99)
    for j in jobs + [extra]:
        srv.submit(j)
    srv.admit_pending()
    srv.step()
    free_before = len(srv.free_slots)
    assert srv.cancel(extra)
    assert extra.result["error"] == "cancelled"
    assert len(srv.free_slots) == free_before + 1
    assert srv.cancel(jobs[1], requeue=True)
    srv.watchdog_ms = 0.0          # the next slice counts as timed out
    srv.admit_pending()
    srv.step()
    assert srv.timeouts == 1 and not srv.active
    srv.watchdog_ms = None
    srv.drain()
    assert len(srv.free_slots) == kw["n_slots"]
    srv.close()
    for a, b in zip(ref, jobs):
        assert _result(a) == _result(b)


@pytest.mark.parametrize("rates", ["f64", "f32"])
def test_large_instance_kernel_matches_small(rates):
    """A pool sized for up to 64 carriers and 7 slots launches the kernel's
    large instance (8 processes per lane, 8 slots per carrier); the same jobs
    in a pool sized for 16 carriers and 4 slots launch the small one.  Every
    result is the same."""
    rng = np.random.default_rng(3)
    specs = [dict(edge=int(rng.integers(3, 5)), n_carriers=int(rng.integers(1, 17)),
                  hops=int(rng.integers(500, 3000)), seed=300 + k,
                  temp=float(rng.uniform(250, 400)),
                  field_vpm=float(rng.choice([0.0, 1e6, 5e6])))
             for k in range(24)]
    out = []
    for nc_max, z_max in ((16, 4), (64, 7)):
        srv = KmcServer(corundum_family, n_slots=32, nc_max=nc_max,
                        z_max=z_max, rows_capacity=1 << 22, slice_hops=700,
                        rate_precision=rates, path_rows=50, msd_lags=20)
        jobs = [Job(k, **s) for k, s in enumerate(specs)]
        for j in jobs:
            srv.submit(j)
        srv.drain()
        srv.close()
        out.append([_result(j) for j in jobs])
    assert all(r[2] > 0 for r in out[0])
    assert out[0] == out[1]


def test_served_jobs_equal_the_batched_trajectories():
    """A job served in slices through the pool gives, bit for bit, the
    result of the same trajectory run in one batched launch."""
    from pycd_amd import batch as B, constants
    from pycd_amd.rng import seeded_starts
    rng = np.random.default_rng(9)
    specs = [dict(edge=int(rng.integers(3, 5)), n_carriers=int(rng.integers(1, 13)),
                  hops=int(rng.integers(800, 2500)), seed=500 + k,
                  temp=float(rng.uniform(260, 380)),
                  field_vpm=float(rng.choice([0.0, 2e6])))
             for k in range(10)]
    srv = KmcServer(corundum_family, n_slots=16, nc_max=16, rows_capacity=1 << 22,
                    slice_hops=311, rate_precision="f64", path_rows=50,
                    msd_lags=20)
    jobs = [Job(k, **s) for k, s in enumerate(specs)]
    for j in jobs:
        srv.submit(j)
    srv.drain()
    n_path = srv.n_path
    srv.close()
    edges = sorted({s["edge"] for s in specs})
    fams = [B.Family(*corundum_family(e)) for e in edges]
    trajs, occs, words = [], [], []
    for s in specs:
        fam = fams[edges.index(s["edge"])]
        o, w = seeded_starts(s["seed"], 1, [fam.hop_sites], [s["n_carriers"]])
        kT = s["temp"] * constants.K2AUTEMP
        field = (np.array([0.0, 0.0, s["field_vpm"] * constants.VPM2AUFIELD])
                 if s["field_vpm"] else np.zeros(3))
        rate = s["n_carriers"] * B.expected_rate_per_carrier(fam, kT, field)
        trajs.append(dict(family=edges.index(s["edge"]),
                          n_carriers=s["n_carriers"], carrier_q=-1.0,
                          field=field, kT=kT,
                          time_interval=s["hops"] / rate / 50,
                          n_path=n_path, max_hops=s["hops"]))
        occs.append(o[0])
        words.append(w[0])
    bt = B.KmcBatch(fams, "f64", "f64")
    bt.setup(trajs, occs, np.asarray(words), waves_per_cu=4)
    bt.run()
    res = bt.results()
    for k, j in enumerate(jobs):
        r = j.result
        assert (r["t"], r["energy"], r["hops"], r["start"]) == \
            (res["t"][k], res["energy"][k], res["hops"][k], res["start"][k]), k
        assert np.array_equal(r["rows"], bt.rows(k)), k
    bt.close()


_MONO = {}


def monoclinic_family(edge):
    if edge not in _MONO:
        w = workloads.monoclinic(edge)
        table = w.build_table()
        _MONO[edge] = (w.make_run(table), table)
    return _MONO[edge]


def test_served_monoclinic_jobs_vs_oracle():
    """The serving form's own lattice (monoclinic, 4 sites per cell, hops to
    2.886 and 2.902 A) through the server in 257-hop slices with FP64
    rates: every job's hop count, clock, path rows and energy equal the
    oracle's trajectory on the same seed and time grid, run on the table
    entries the served family reads (its cell-0 rows)
    (PyCD/core.py:2647-2861)."""
    import random
    from oracle import kmc, seeding
    from pycd_amd import constants
    from tests.systems import cell0_table, oracle_problem
    rng = np.random.default_rng(21)
    specs = [dict(edge=int(rng.integers(4, 7)),
                  n_carriers=int(rng.integers(1, 9)),
                  hops=int(rng.integers(600, 1500)), seed=700 + k,
                  temp=float(rng.uniform(260, 380)),
                  field_vpm=float(rng.choice([0.0, 3e6])))
             for k in range(8)]
    srv = KmcServer(monoclinic_family, n_slots=16, nc_max=8, z_max=2,
                    rows_capacity=1 << 22, slice_hops=257,
                    rate_precision="f64", path_rows=40, msd_lags=10)
    jobs = [Job(k, **s) for k, s in enumerate(specs)]
    for j in jobs:
        srv.submit(j)
    srv.drain()
    n_path = srv.n_path
    srv.close()
    assert len({s["edge"] for s in specs}) > 1
    cell0 = {}
    for j, s in zip(jobs, specs):
        run, table = monoclinic_family(s["edge"])
        if s["edge"] not in cell0:
            cell0[s["edge"]] = cell0_table(
                table, run.system_size,
                table.shape[0] // int(np.prod(run.system_size)))
        table = cell0[s["edge"]]
        nc = s["n_carriers"]
        field = (np.array([0.0, 0.0, 1.0]) * s["field_vpm"]
                 * constants.VPM2AUFIELD if s["field_vpm"] else np.zeros(3))
        pb = oracle_problem(run, table=table, field=field)
        pb.kT = s["temp"] * constants.K2AUTEMP
        pb.carrier_q = np.full(nc, -1.0)
        pb.time_interval = j.time_interval
        pb.n_path = n_path
        gen = random.Random()
        gen.setstate(seeding.trajectory_states(s["seed"], 1)[0])
        occ = seeding.initial_occupancy(gen, run.hop_site_lists["electron"],
                                        nc)
        ora = kmc.run(pb, occ, gen.random, max_hops=s["hops"])
        r = j.result
        assert r["hops"] == ora["hops"] > 0, j.id
        assert abs(r["t"] - ora["time_data"][-1]) <= \
            1e-12 * ora["time_data"][-1], j.id
        assert r["start"] == ora["start"], j.id
        assert np.max(np.abs(r["rows"] - ora["unwrapped"])) < 1e-8, j.id
        scale = max(1.0, abs(ora["energy_now"]))
        assert abs(r["energy"] - ora["energy_now"]) < 1e-9 * scale, j.id

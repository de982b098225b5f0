"""Host-side logic that needs no GPU: queue order, row arena, job streams,
and the numbers the benchmark line takes from the committed profiles."""
from pathlib import Path

import numpy as np

import bench
from pycd_amd import tuning
from pycd_amd.serve import RowArena, poisson_jobs

ROOT = Path(__file__).resolve().parents[1]


class _Fam:
    def __init__(self, z):
        self.z = z


def test_lpt_order_is_a_stable_descending_cost_permutation():
    fams = [_Fam(4), _Fam(6)]
    rs = np.random.default_rng(0)
    trajs = [dict(family=int(rs.integers(0, 2)),
                  n_carriers=int(rs.integers(1, 65)),
                  max_hops=int(rs.choice([1000, 100000])))
             for _ in range(500)]
    cost = tuning.expected_cost(trajs, fams)
    order = tuning.lpt_order(trajs, fams)
    assert sorted(order.tolist()) == list(range(len(trajs)))
    c = cost[order]
    assert np.all(c[:-1] >= c[1:])
    # equal costs keep their input order
    for a, b in zip(order[:-1], order[1:]):
        if cost[a] == cost[b]:
            assert a < b
    # more carriers, more slots or more hops never cost less
    t = dict(family=0, n_carriers=10, max_hops=1000)
    base = tuning.expected_cost([t], fams)[0]
    for more in (dict(n_carriers=11), dict(family=1), dict(max_hops=2000)):
        assert tuning.expected_cost([{**t, **more}], fams)[0] > base


def test_row_arena_first_fit_and_merge():
    a = RowArena(1000)
    blocks = [a.alloc(n) for n in (100, 200, 300, 400)]
    assert blocks == [0, 100, 300, 600]
    assert a.alloc(1) is None and a.used == 1000
    a.release(100, 200)
    a.release(600, 400)
    assert a.alloc(250) == 600          # first block large enough
    assert a.alloc(150) == 100
    a.release(0, 100)
    a.release(100, 150)
    a.release(300, 300)
    a.release(600, 250)
    assert a.free == [(0, 1000)] and a.used == 0


def test_poisson_jobs_are_reproducible_and_in_range():
    j1 = poisson_jobs(200, rate=200.0, edges=(4, 12), nc=(1, 32),
                      hops=(1e5, 1e6), seed=5)
    j2 = poisson_jobs(200, rate=200.0, edges=(4, 12), nc=(1, 32),
                      hops=(1e5, 1e6), seed=5)
    assert [(j.edge, j.n_carriers, j.hops, j.arrival) for j in j1] == \
        [(j.edge, j.n_carriers, j.hops, j.arrival) for j in j2]
    arr = np.array([j.arrival for j in j1])
    assert np.all(np.diff(arr) >= 0)
    # mean inter-arrival time of a 200/s Poisson stream, within 4 sigma
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 1 / 200) < 4 * (1 / 200) / np.sqrt(len(gaps))
    assert all(4 <= j.edge <= 12 and 1 <= j.n_carriers <= 32 and
               1e5 <= j.hops <= 1e6 for j in j1)


def test_bench_reads_the_committed_traffic_profile():
    t = bench.measured_traffic(ROOT / "profiles" / "batched_kmc_4096traj.md")
    assert t is not None and 1e9 < t < 1e12
    assert bench.measured_traffic(ROOT / "profiles" / "missing.md") is None


def test_bench_refuses_more_than_one_gpu(capsys, monkeypatch):
    monkeypatch.setattr("sys.argv", ["bench.py", "--gpus", "2"])
    try:
        bench.main()
    except SystemExit as e:
        assert e.code == 2
    else:
        raise AssertionError("bench.py accepted --gpus 2")
    assert "one GPU" in capsys.readouterr().err


class _StalledServer:
    """KmcServer whose every slice overruns the watchdog; admission and slot
    release are stubbed, so no device is touched."""

    @staticmethod
    def make(**kw):
        from pycd_amd.serve import KmcServer

        class Stalled(KmcServer):
            launches = 0

            def _try_admit(self, job):
                if not self.free_slots:
                    return False
                s = self.free_slots.popleft()
                job.slot, job.rows_at, job.n_path = s, 0, 1
                self.active[s] = job
                return True

            def _never_fits(self, job):
                return None

            def _release_slot(self, slot):
                pass

            def _launch_slice(self):
                Stalled.launches += 1
                return 1e9

        srv = Stalled(lambda edge: None, n_slots=4, rows_capacity=1 << 10,
                      watchdog_ms=10.0, **kw)
        srv.rows.release = lambda lo, n: None
        return srv


def test_watchdog_requeue_is_bounded():
    """Slices that always overrun the watchdog: each job is re-queued
    max_timeouts times and then rejected, so drain() ends."""
    from pycd_amd.serve import Job
    srv = _StalledServer.make(max_timeouts=2)
    jobs = [Job(k, edge=3, n_carriers=1, hops=10, seed=k) for k in range(6)]
    for j in jobs:
        srv.submit(j)
    srv.drain()
    assert all(j.result["error"] == "timed out in 3 slices" for j in jobs)
    assert srv.requeued == 2 * len(jobs)
    assert not srv.active and not srv.pending
    srv.h = None


def test_footprint_is_measured_once_per_edge(monkeypatch):
    """The table footprint of an edge is computed from one family build,
    which the following admission reuses instead of building again."""
    from pycd_amd import serve
    calls = []

    class FakeFamily:
        def __init__(self, run, table):
            calls.append(run)

        def device_bytes(self, precision):
            return 1234

    monkeypatch.setattr(serve, "Family", FakeFamily)
    srv = serve.KmcServer(lambda edge: (edge, None), n_slots=2)
    assert srv.table_bytes(5) == 1234
    assert srv.table_bytes(5) == 1234
    assert calls == [5]
    assert srv._staged[0] == 5
    srv.h = None

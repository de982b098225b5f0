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

"""Batched persistent KMC kernel vs the oracle, the MSD oracle and itself."""
import random

import numpy as np
import pytest

from oracle import kmc, msd as omsd, seeding
from pycd_amd import batch as B, constants, rng, tuning, workloads
from tests.systems import cell0_table, oracle_problem

pytestmark = pytest.mark.gpu


_CELL0 = {}


def _families(sizes, builder=workloads.corundum, **kw):
    _CELL0.clear()
    fams, runs = [], []
    for n in sizes:
        w = builder(n, **kw)
        table = w.build_table()
        run = w.make_run(table)
        fams.append(B.Family(run, table))
        runs.append((w, run, table))
    return fams, runs


def _mixed_batch(fams, runs, specs, seed, max_hops):
    """specs: list of (family, n_carriers, field_vpm, T)."""
    n_traj = len(specs)
    sites = [fams[f].hop_sites for f, _, _, _ in specs]
    counts = [nc for _, nc, _, _ in specs]
    occs, words = rng.seeded_starts(seed, n_traj, sites, counts)
    trajs = []
    for (f, nc, fv, T) in specs:
        field = np.array([0.0, 0.0, fv * constants.VPM2AUFIELD])
        trajs.append(dict(family=f, n_carriers=nc, carrier_q=-1.0,
                          field=field, kT=T * constants.K2AUTEMP,
                          time_interval=runs[f][1].time_interval,
                          n_path=int(runs[f][1].t_final
                                     / runs[f][1].time_interval) + 1,
                          max_hops=max_hops))
    return trajs, occs, words


def _oracle(runs, spec, traj, seed, t_index, n_traj, max_hops):
    """The oracle's trajectory on the table entries the family holds (its
    cell-0 rows, moved; equal to the full table to rounding).  Rates move by
    a rounding difference in dG0 over kT, so at kT ~ 1e-3 Ha ulp-level
    table differences alone can move a clock by ~1e-12."""
    f, nc, fv, T = spec
    w, run, table = runs[f]
    if f not in _CELL0 or _CELL0[f][0] is not table:
        n_unit = table.shape[0] // int(np.prod(run.system_size))
        _CELL0[f] = (table, cell0_table(table, run.system_size, n_unit))
    pb = oracle_problem(run, table=_CELL0[f][1],
                        field=np.asarray(traj["field"]))
    pb.kT = traj["kT"]
    pb.carrier_q = np.full(nc, -1.0)
    state = seeding.trajectory_states(seed, n_traj)[t_index]
    gen = random.Random()
    gen.setstate(state)
    occ = seeding.initial_occupancy(gen, run.hop_site_lists["electron"], nc)
    return kmc.run(pb, occ, gen.random, max_hops=max_hops)


def test_fp64_batch_matches_oracle_per_trajectory():
    fams, runs = _families([3, 4], t_final=2e-8)
    specs = [(0, 1, 0.0, 300.0), (1, 5, 1e6, 300.0), (0, 16, 5e6, 350.0),
             (1, 3, 0.0, 250.0), (1, 16, 2e6, 400.0), (0, 2, 1e7, 280.0)]
    _fp64_vs_oracle(fams, runs, specs, 7, 1500)


def test_fp64_batch_matches_oracle_many_carriers():
    """The batched form's large end: 32 and 64 electrons on the 6x6x6
    (2592 sites) and 8x8x8 (6144 sites) families, fields up to 1e7 V/m,
    250-400 K; per trajectory the oracle's hops, clock, rows, energy and
    drift (the kernel's large instance runs the 64-electron ones)."""
    fams, runs = _families([6, 8], t_final=2e-8)
    specs = [(0, 32, 1e6, 300.0), (1, 64, 5e6, 350.0), (1, 32, 0.0, 250.0),
             (0, 64, 1e7, 400.0)]
    _fp64_vs_oracle(fams, runs, specs, 13, 800)


def _fp64_vs_oracle(fams, runs, specs, seed, max_hops):
    trajs, occs, words = _mixed_batch(fams, runs, specs, seed, max_hops)
    bt = B.KmcBatch(fams, "f64", "f64")
    order = np.argsort([-s[1] for s in specs], kind="stable")
    bt.setup(trajs, occs, words, order=order, waves_per_cu=2)
    bt.run()
    res = bt.results()
    for k, spec in enumerate(specs):
        ora = _oracle(runs, spec, trajs[k], seed, k, len(specs), max_hops)
        assert res["hops"][k] == ora["hops"] == max_hops
        assert abs(res["t"][k] - ora["time_data"][-1]) <= \
            1e-12 * ora["time_data"][-1], k
        rows = bt.rows(k)
        assert np.max(np.abs(rows - ora["unwrapped"])) < 1e-8, k
        scale = max(1.0, abs(ora["energy_now"]))
        assert abs(res["energy"][k] - ora["energy_now"]) < 1e-9 * scale, k
        assert res["start"][k] == ora["start"]
        nc = spec[1]
        # the batched family uses cell 0's hop vectors for every site (equal
        # to each site's own to rounding), so a component that cancels to
        # ~1e-17 is compared against the drift's magnitude
        dscale = np.max(np.abs(ora["drift"]))
        assert np.allclose(res["drift"][k, :nc], ora["drift"], rtol=1e-9,
                           atol=1e-12 * dscale)
    bt.close()


def test_msd_and_graph_replays():
    fams, runs = _families([3], t_final=2e-10, time_interval=1e-13)
    specs = [(0, nc, 0.0, 300.0) for nc in (1, 4, 9, 12)]
    seed = 11
    trajs, occs, words = _mixed_batch(fams, runs, specs, seed, 4000)
    lags = [200, 150, 100, 64]
    scale = 1 / constants.ANG2BOHR
    bt = B.KmcBatch(fams, "f64", "f32")
    bt.setup(trajs, occs, words, msd_lags=lags, msd_scale=scale,
             waves_per_cu=1)
    bt.run(use_graph=False)
    eager = bt.results()
    eager_msd = bt.msd()
    for k in range(len(specs)):
        rows = bt.rows(k)
        P = int(eager["start"][k])
        nc = specs[k][1]
        pos = rows[:P].reshape(P, nc, 3) * scale
        ref = omsd.squared_displacement(pos, 1, P, lags[k])[0].mean(axis=1)
        assert np.allclose(eager_msd[k], ref, rtol=1e-11, atol=1e-300)
    # graph replays with eager runs in between reproduce the eager results
    for _ in range(3):
        bt.run(use_graph=True)
        rep = bt.results()
        for key in ("t", "hops", "energy", "start"):
            assert np.array_equal(rep[key], eager[key]), key
        for a, b in zip(bt.msd(), eager_msd):
            assert np.array_equal(a, b)
        bt.run(use_graph=False)
    bt.close()


def test_queue_bins_leave_results_unchanged():
    """Per-SIMD bins (16 waves per CU), per-CU bins (2) and one queue (1)
    hand trajectories to different wavefronts; every result is the same."""
    fams, runs = _families([3], t_final=2e-8)
    rs = np.random.default_rng(5)
    specs = [(0, int(nc), float(fv), 300.0)
             for nc, fv in zip(rs.integers(1, 9, 1100),
                               rs.choice([0.0, 1e6, 5e6], 1100))]
    trajs, occs, words = _mixed_batch(fams, runs, specs, 17, 120)
    order = np.argsort([-s[1] for s in specs], kind="stable")
    out = []
    for wpc in (16, 2, 1):
        bt = B.KmcBatch(fams, "f64", "f32")
        bt.setup(trajs, occs, words, order=order, waves_per_cu=wpc)
        bt.run()
        res = bt.results()
        out.append((res, [bt.rows(k) for k in (0, 1, 777, 1099)]))
        bt.close()
    ref, ref_rows = out[0]
    # every trajectory ran to one of its limits: max_hops or a full path
    n_path = np.array([t["n_path"] for t in trajs])
    assert np.all((ref["hops"] == 120) | (ref["start"] == n_path))
    assert np.all(ref["hops"] > 0)
    for res, rows in out[1:]:
        for key in ("t", "hops", "energy", "start", "drift"):
            assert np.array_equal(res[key], ref[key]), key
        for a, b in zip(rows, ref_rows):
            assert np.array_equal(a, b)


def test_handover_leaves_results_unchanged():
    """Short and long trajectories on per-SIMD bins: wavefronts that run out
    of work take the long ones over at hop boundaries, from their saved
    state.  Every result equals the run without hand-over, bit for bit,
    eager and replayed from the graph, and hand-overs did happen."""
    fams, runs = _families([3], t_final=2e-7, time_interval=1e-11)
    rs = np.random.default_rng(9)
    n = 1100
    specs = [(0, int(nc), float(fv), 300.0)
             for nc, fv in zip(rs.integers(1, 9, n),
                               rs.choice([0.0, 1e6], n))]
    trajs, occs, words = _mixed_batch(fams, runs, specs, 23, 300)
    long = rs.choice(n, 120, replace=False)
    for k in long:
        trajs[k]["max_hops"] = 12000
    order = tuning.lpt_order(trajs, fams)
    out = {}
    for hand in (False, True):
        bt = B.KmcBatch(fams, "f64", "f32")
        bt.setup(trajs, occs, words, order=order, waves_per_cu=4)
        bt.set_handover(hand)
        bt.run()
        res = bt.results()
        out[hand] = (res, [bt.rows(int(k)) for k in long[:6]],
                     bt.handovers())
        if hand:
            bt.run(use_graph=True)
            bt.run(use_graph=True)
            out["graph"] = (bt.results(), [bt.rows(int(k)) for k in long[:6]],
                            bt.handovers())
        bt.close()
    ref, ref_rows, none = out[False]
    assert none == 0
    # every trajectory ran to one of its limits: max_hops or a full path
    n_path = np.array([t["n_path"] for t in trajs])
    max_hops = np.array([t["max_hops"] for t in trajs])
    assert np.all((ref["hops"] == max_hops) | (ref["start"] == n_path))
    assert np.mean(ref["hops"][long] == 12000) > 0.9
    print("hand-overs:", out[True][2], out["graph"][2])
    for key in (True, "graph"):
        res, rows, count = out[key]
        assert count > 0
        for f in ("t", "hops", "energy", "start", "drift", "no_select"):
            assert np.array_equal(res[f], ref[f]), (key, f)
        for a, b in zip(rows, ref_rows):
            assert np.array_equal(a, b)


def test_hop_differences_leave_results_unchanged():
    """With its families' hop differences a launch gathers each slot's
    potential change A[n, x] - A[s, x] as one entry; without them (switch
    off, or a family too large to hold them) it gathers both entries.  The
    numbers are the same: every result equal bit for bit, for FP64 and FP32
    rates and the FP32 table, on corundum and monoclinic pairs of families
    in one launch (1100 trajectories over per-SIMD bins, so the XCD shares
    of the queue hold different families)."""
    for builder, sizes in ((workloads.corundum, [3, 4]),
                           (workloads.monoclinic, [4, 5])):
        fams, runs = _families(sizes, builder=builder, t_final=2e-8)
        rs = np.random.default_rng(31)
        n = 1100
        specs = [(int(f), int(nc), float(fv), float(T))
                 for f, nc, fv, T in zip(rs.integers(0, 2, n),
                                         rs.integers(1, 17, n),
                                         rs.choice([0.0, 1e6, 5e6], n),
                                         rs.uniform(250.0, 400.0, n))]
        trajs, occs, words = _mixed_batch(fams, runs, specs, 41, 300)
        order = tuning.lpt_order(trajs, fams)
        for table, rates in (("f64", "f64"), ("f64", "f32"), ("f32", "f32")):
            out = []
            for on in (True, False):
                bt = B.KmcBatch(fams, table, rates)
                bt.set_hop_differences(on)
                bt.setup(trajs, occs, words, order=order, waves_per_cu=16)
                bt.run()
                out.append((bt.results(),
                            [bt.rows(k) for k in (0, 1, 550, n - 1)]))
                bt.close()
            (ra, rows_a), (rb, rows_b) = out
            assert np.all(ra["hops"] > 0)
            for key in ("t", "hops", "energy", "start", "drift", "no_select"):
                assert np.array_equal(ra[key], rb[key]), (table, rates, key)
            for a, b in zip(rows_a, rows_b):
                assert np.array_equal(a, b)


def test_fp64_batch_monoclinic_matches_oracle_per_trajectory():
    """The serving form's monoclinic lattice (C-centred, 4 sites per cell,
    two hop slots) in the batched FP64 mode, against the oracle."""
    fams, runs = _families([4, 5], builder=workloads.monoclinic,
                           t_final=2e-8)
    specs = [(0, 1, 0.0, 300.0), (1, 7, 1e6, 320.0), (0, 12, 5e6, 280.0),
             (1, 3, 0.0, 390.0)]
    seed, max_hops = 13, 1200
    trajs, occs, words = _mixed_batch(fams, runs, specs, seed, max_hops)
    bt = B.KmcBatch(fams, "f64", "f64")
    bt.setup(trajs, occs, words, waves_per_cu=2)
    bt.run()
    res = bt.results()
    for k, spec in enumerate(specs):
        ora = _oracle(runs, spec, trajs[k], seed, k, len(specs), max_hops)
        assert res["hops"][k] == ora["hops"] == max_hops
        assert abs(res["t"][k] - ora["time_data"][-1]) <= \
            1e-12 * ora["time_data"][-1], k
        assert np.max(np.abs(bt.rows(k) - ora["unwrapped"])) < 1e-8, k
        scale = max(1.0, abs(ora["energy_now"]))
        assert abs(res["energy"][k] - ora["energy_now"]) < 1e-9 * scale, k
        assert res["start"][k] == ora["start"]
    bt.close()


def test_msd_lag_passes_and_many_carriers():
    """The MSD kernels take lags in passes of 512 (two blocks of 256) and
    one workgroup per (trajectory, carrier): lags reaching into the second
    and third pass, and 64, 33 and 2 electrons, against the oracle's squared
    displacements (each path filled: 2001 rows)."""
    fams, runs = _families([4], t_final=2e-8)
    specs = [(0, 64, 0.0, 300.0), (0, 2, 1e6, 350.0), (0, 33, 0.0, 280.0)]
    seed = 23
    trajs, occs, words = _mixed_batch(fams, runs, specs, seed, 60000)
    for t, (f, nc, fv, T) in zip(trajs, specs):
        rate = B.expected_rate_per_carrier(fams[f], t["kT"], t["field"]) * nc
        t["time_interval"] = 6000.0 / rate / 2000
        t["n_path"] = 2001
    lags = [1300, 700, 577]
    scale = 1 / constants.ANG2BOHR
    bt = B.KmcBatch(fams, "f64", "f32")
    bt.setup(trajs, occs, words, msd_lags=lags, msd_scale=scale,
             waves_per_cu=1)
    bt.run(use_graph=False)
    res = bt.results()
    got = bt.msd()
    for k, (f, nc, fv, T) in enumerate(specs):
        P = int(res["start"][k])
        assert P == 2001, (k, P, res["hops"][k])
        pos = bt.rows(k)[:P].reshape(P, nc, 3) * scale
        ref = omsd.squared_displacement(pos, 1, P, lags[k])[0].mean(axis=1)
        assert got[k][0] == 0.0
        assert np.allclose(got[k], ref, rtol=1e-11, atol=1e-300), k
    bt.close()


def test_msd_partly_written_paths():
    """Paths the runs fill only in part (151, 420, 918, 3544 and 312 of
    4001 rows: the rate estimate that sets the time grid counts no blocked
    hops, so the 8-electron run writes more rows than its target), with lags
    past the written rows, one and three passes of lags and 1-8 electrons:
    the lags below the written rows against the oracle's squared
    displacements, the rest zero."""
    fams, runs = _families([4], t_final=2e-8)
    specs = [(0, 1, 0.0, 300.0), (0, 3, 1e6, 350.0), (0, 5, 0.0, 280.0),
             (0, 8, 0.0, 320.0), (0, 2, 0.0, 300.0)]
    targets = [150, 400, 700, 1500, 300]
    lags = [500, 500, 500, 1100, 1200]
    hops = 6000
    trajs, occs, words = _mixed_batch(fams, runs, specs, 29, hops)
    for t, (f, nc, fv, T), p in zip(trajs, specs, targets):
        rate = B.expected_rate_per_carrier(fams[f], t["kT"], t["field"]) * nc
        t["time_interval"] = hops / rate / p
        t["n_path"] = 4001
    scale = 1 / constants.ANG2BOHR
    bt = B.KmcBatch(fams, "f64", "f32")
    bt.setup(trajs, occs, words, msd_lags=lags, msd_scale=scale,
             waves_per_cu=1)
    bt.run(use_graph=False)
    res = bt.results()
    got = bt.msd()
    seen = []
    for k, (f, nc, fv, T) in enumerate(specs):
        P = int(res["start"][k])
        assert 1 < P < 4001, (k, P, list(res["start"]))
        seen.append(P)
        m = min(P, lags[k])
        pos = bt.rows(k)[:P].reshape(P, nc, 3) * scale
        ref = omsd.squared_displacement(pos, 1, P, m)[0].mean(axis=1)
        assert got[k][0] == 0.0
        assert np.allclose(got[k][:m], ref, rtol=1e-11, atol=1e-300), k
        assert np.all(got[k][m:] == 0.0), k
    # the written rows fall below one block of lags, inside the second block
    # and past the pass
    assert min(seen) < 256 and any(256 < P < 512 for P in seen)
    assert max(seen) > 1100
    bt.close()

"""Schedules and tile shapes, chosen by measurement.

`autotune_table` times the table kernels' tile variants for one
precision on the caller's own lattice (the table build is timed on the
device) and keeps the fastest in the context; every variant gives the same
table bit for bit.

Two knobs decide how trajectories map onto the chip in the batched kernel:
  * the launch order of the work queue -- longest expected work first
    (LPT).  The library deals this order out to one bin per SIMD in
    serpentine order, so each SIMD gets a mix of long and short
    trajectories, and the wavefronts sharing a SIMD pace each other;
  * waves per CU of the persistent grid -- more waves hide more gather
    latency but split the CU's issue slots and LDS.
`autotune_waves` times short runs of the real batch at each candidate and
keeps the fastest.
"""
import numpy as np


def expected_cost(trajs, fams):
    """Relative per-trajectory work: hops x (fixed part + per-carrier part)."""
    cost = np.empty(len(trajs))
    for k, t in enumerate(trajs):
        z = fams[t["family"]].z
        nc = t["n_carriers"]
        hops = t.get("max_hops", 0) or 1
        cost[k] = hops * (8.0 + 3 * (1 + z) * nc / 64.0 * 4 + nc * z / 64.0)
    return cost


def lpt_order(trajs, fams):
    """Queue order: descending expected cost (stable)."""
    return np.argsort(-expected_cost(trajs, fams), kind="stable").astype(
        np.int32)


def autotune_waves(batch, trajs, occs, words, order, candidates=(2, 4, 6, 8,
                                                                  12, 16),
                   sample=4096, hops=2000, log=None):
    """Time `hops`-hop runs of the first `sample` queued trajectories.

    The sample must hold several trajectories per wavefront at the largest
    candidate, or every candidate past n_traj / CUs times one trajectory
    per wave and the comparison says nothing; the default takes the whole
    batch of the benchmark workload, each cut to `hops` hops."""
    sub = [int(i) for i in order[:sample]]
    sub_trajs = []
    for i in sub:
        t = dict(trajs[i])
        t["max_hops"] = hops
        sub_trajs.append(t)
    sub_occs = [occs[i] for i in sub]
    sub_words = np.asarray(words)[sub]
    sub_order = np.arange(len(sub), dtype=np.int32)
    results = {}
    for w in candidates:
        batch.setup(sub_trajs, sub_occs, sub_words, order=sub_order,
                    waves_per_cu=w)
        batch.run()
        ms = min(batch.run() for _ in range(2))
        h = batch.results()["hops"].sum()
        results[w] = h / (ms / 1e3)
        if log:
            log(f"  autotune waves/CU={w}: {results[w]:.4e} hops/s")
    best = max(results, key=results.get)
    return best, results


def autotune_table(pos, cell, trans, pbc, alpha, r_cut, k_cut, eps_r,
                   precision="f64", repeats=2, log=None):
    """Build the table once per tile variant of `precision` (after one
    warm-up build), keep the variant whose device time (structure factors,
    contraction and epilogue) is shortest, and set it for later
    `pair_table` calls.  Returns (variant, {variant: ms})."""
    from . import _native
    times = {}
    out = None
    for v in _native.TABLE_VARIANTS[precision]:
        _native.set_table_variant(precision, v)
        best = float("inf")
        for r in range(repeats + 1):
            out, rep = _native.pair_table(pos, cell, trans, pbc, alpha, r_cut,
                                          k_cut, eps_r, precision, out=out)
            if r > 0:  # the first build of a variant warms it up
                best = min(best, rep["ms_total"])
        times[v] = best
        if log:
            log(f"  autotune table {precision} variant {v}: {best:.2f} ms")
    chosen = min(times, key=times.get)
    _native.set_table_variant(precision, chosen)
    return chosen, times

"""Schedules for the batched kernel, chosen by measurement.

Two knobs decide how trajectories map onto the chip:
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

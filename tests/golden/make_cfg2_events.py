"""Writes tests/golden/cfg2_events.npz: the oracle's 1e6-hop event record of
the 8x8x8 corundum configuration (6144 sites, 16 electrons, 1e6 V/m along
c, seed 2), against which the plain GPU path is compared hop for hop.

    python tests/golden/make_cfg2_events.py [--hops 1000000] [--workers 8]

Everything here runs on the CPU with the oracle: hop lists from
oracle.neighbors, the FP64 pair table row by row from oracle.ewald
(in parallel), and the trajectory from oracle.kmc with the reference's
seeding protocol (oracle.seeding).  The full record is 4 MB of process
indices and 170 MB of path arrays, so the fixture keeps:

  block_hash   BLAKE2b-64 of each 1000-hop block of chosen process indices
               (uint8), so the whole event sequence is compared bit for bit
  time_k       the clock after every 1000th hop (hop 0, 1000, ...)
  rows_k, energy_k, delg_k   every 100th path row of unwrapped_traj,
               energy_traj and delG0_traj, and row_index
  occupancy, drift, energy_now, t, start, hops   the final state
"""
import argparse
import hashlib
import multiprocessing as mp
import random
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

BLOCK = 1000
ROW_STRIDE = 100


def block_hashes(procs, block=BLOCK):
    """BLAKE2b-64 digest of each `block` hops of process indices as uint8
    (the plain path has at most 64 processes here)."""
    p = np.asarray(procs)
    assert p.max() < 256
    p = p.astype(np.uint8)
    out = []
    for k in range(0, len(p), block):
        h = hashlib.blake2b(p[k:k + block].tobytes(), digest_size=8)
        out.append(int.from_bytes(h.digest(), "little"))
    return np.asarray(out, dtype=np.uint64)


def workload():
    from oracle.neighbors import hop_neighbor_list
    from pycd_amd import workloads
    return workloads.corundum(8, n_carriers=16, field_vpm=1e6, t_final=2e-8,
                              hop_lists=hop_neighbor_list)


_W = None


def _rows(lo_hi):
    from oracle import ewald
    w = _W
    s = w.system
    return ewald.table_rows(w.neighbors.bulk_sites.cell_coordinates,
                            range(*lo_hi), w.material.lattice_matrix, w.size,
                            w.neighbors.pbc, s.alpha, s.r_cut, s.k_cut,
                            w.material.dielectric_constant)


def oracle_table(w, workers):
    global _W
    _W = w
    n = w.n_sites
    step = 96
    spans = [(k, min(k + step, n)) for k in range(0, n, step)]
    table = np.empty((n, n))
    with mp.get_context("fork").Pool(workers) as pool:
        for (lo, hi), rows in zip(spans, pool.imap(_rows, spans)):
            table[lo:hi] = rows
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--hops", type=int, default=1_000_000)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=str(Path(__file__).with_name(
        "cfg2_events.npz")))
    args = ap.parse_args()
    from oracle import kmc, seeding
    from tests.systems import oracle_problem
    t0 = time.time()
    w = workload()
    table = oracle_table(w, args.workers)
    print(f"table {w.n_sites}^2 in {time.time() - t0:.0f} s", flush=True)
    run = w.make_run(table)
    state = seeding.trajectory_states(2, 1)[0]
    gen = random.Random()
    gen.setstate(state)
    occ = seeding.initial_occupancy(gen, run.hop_site_lists["electron"], 16)
    t0 = time.time()
    res = kmc.run(oracle_problem(run, table=table), occ, gen.random,
                  max_hops=args.hops, record=True)
    print(f"{res['hops']} hops in {time.time() - t0:.0f} s", flush=True)
    assert res["hops"] == args.hops, "path filled before the hop budget"
    idx = np.arange(0, len(res["energy"]), ROW_STRIDE)
    np.savez_compressed(
        args.out, block=BLOCK, block_hash=block_hashes(res["procs"]),
        time_k=res["time_data"][::BLOCK], row_index=idx,
        rows_k=res["unwrapped"][idx], energy_k=res["energy"][idx],
        delg_k=res["delg_0"][idx], occupancy=np.asarray(res["occupancy"]),
        initial_occupancy=np.asarray(occ), drift=res["drift"],
        energy_now=res["energy_now"], t=res["time_data"][-1],
        start=res["start"], hops=res["hops"])
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()

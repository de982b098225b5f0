"""Benchmark of the KMC hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode batched|plain]

Headline workload (`batched`): 4096 independent trajectories in one
persistent launch -- Fe-only corundum supercells with edge U{4..8} (768 to
6144 sites), 1..64 electrons, |E| ~ U(0, 1e7) V/m along a random lattice
direction, T ~ U(250, 400) K, 1e5 hops each, FP32 rates with an FP64
prefix sum and clock, per-trajectory MSD reduced on the device.  One step
= one pass over the whole batch (4.1e8 hops).  Inputs (tables, hop slots,
generator states, initial occupancies) are resident in HBM before timing.

Beside it the plain one-trajectory path is timed on its own workload
(8x8x8 supercell, 6144 sites, 16 electrons, 1e6 V/m along c, 1e6 hops,
FP64, uniforms resident in HBM), and the NumPy oracle on a bounded sample
of the batched workload gives the CPU baseline.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# workloads built by the GPU legs that the CPU legs time the oracle on
_KEEP = {}


def _json_default(o):
    """numpy scalars and arrays in the JSON line"""
    if hasattr(o, "tolist"):
        return o.tolist()
    raise TypeError(f"not JSON serialisable: {type(o).__name__}")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["batched", "precision", "plain"],
                    default="batched")
    ap.add_argument("--precision-edge", type=int, default=10)
    ap.add_argument("--precision-hops", type=int, default=1_000_000)
    ap.add_argument("--n-traj", type=int, default=4096)
    ap.add_argument("--hops", type=int, default=100_000)
    ap.add_argument("--plain-hops", type=int, default=1_000_000)
    ap.add_argument("--plain-variant", choices=["auto", "wave", "workgroup"],
                    default="auto")
    ap.add_argument("--no-precision", action="store_true",
                    help="default run: leave out the precision form")
    ap.add_argument("--waves-per-cu", type=int, default=0,
                    help="0: autotune")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-single", action="store_true",
                    help="CPU baseline on one core only")
    ap.add_argument("--no-handover", action="store_true",
                    help="batched: no hand-over of trajectories to idle "
                         "wavefronts")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=1000)
    ap.add_argument("--serve-slice", type=int, default=4000)
    ap.add_argument("--traffic-profile",
                    default=str(ROOT / "profiles" / "batched_kmc_4096traj.md"),
                    help="profile summary holding the measured HBM bytes per "
                         "launch of the batched kernel")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="bytes of a device-to-device copy issued first "
                         "(counter calibration under rocprofv3 --pmc)")
    return ap.parse_args()


def algorithmic_bytes_batched(trajs, fams, hops, starts, table_bytes,
                              refresh=4096, diffs=True):
    """Bytes the batched algorithm touches per launch.

    Per hop of a trajectory with n_c carriers and z slots:
      slot updates    (1 + z) n_c FP64 hop differences, one per slot row
                      (with diffs=False: 2 (1 + z) n_c table entries)
      fresh sums      (z - 1) n_c table entries (the mover's new slots)
      refresh         (1 + z) n_c^2 / refresh table entries
      hop records     z x 72 B (the new processes) + 8 B (hop vector)
    plus 3 n_c doubles per path row written (the carriers' displacement and
    drift stay in their lanes between path rows).
    """
    total = 0.0
    for k, t in enumerate(trajs):
        z = fams[t["family"]].z
        nc = t["n_carriers"]
        upd = (1 + z) * nc * 8.0 if diffs else 2 * (1 + z) * nc * table_bytes
        entries = (z - 1) * nc + (1 + z) * nc * nc / refresh
        per_hop = upd + entries * table_bytes + z * 72 + 8
        total += per_hop * hops[k] + 24.0 * nc * starts[k]
    return total


def measured_traffic(path):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    import re
    try:
        text = Path(path).read_text()
    except OSError:
        return None
    m = re.search(r"traffic per launch(?: \(grid \d+\))? = .*? = ([0-9.e+]+) B",
                  text)
    return float(m.group(1)) if m else None


def run_batched(args, log):
    from pycd_amd import workloads, batch as B, tuning, constants
    t0 = time.time()
    fams, works, trajs, occs, words, lags = workloads.batched_config(
        n_traj=args.n_traj, hops=args.hops)
    log(f"batched setup: {len(fams)} families, {len(trajs)} trajectories "
        f"({time.time() - t0:.1f}s)")
    order = tuning.lpt_order(trajs, fams)
    wpc = args.waves_per_cu
    bt = B.KmcBatch(fams, table_precision="f64", rate_precision="f32")
    if wpc <= 0:
        wpc, table = tuning.autotune_waves(bt, trajs, occs, words, order,
                                           log=log)
    bt.setup(trajs, occs, words, order=order, msd_lags=lags,
             msd_scale=1 / constants.ANG2BOHR, waves_per_cu=wpc)
    bt.set_handover(not args.no_handover)
    for _ in range(args.warmup):
        bt.run(use_graph=True)
    times = []
    for _ in range(args.steps):
        times.append(bt.run(use_graph=True))
    res = bt.results()
    hops = res["hops"].astype(float)
    total_hops = hops.sum()
    ms = float(np.mean(times))
    value = total_hops / (ms / 1e3)
    # kernel-only timing (no graph) for the roofline of the dominant kernel
    bt.run(use_graph=False)
    bt_k = bt.kernel_ms()
    log(f"  {bt.handovers()} trajectories handed over to idle wavefronts")
    busy = res["busy_s"]
    ncs = np.array([t["n_carriers"] for t in trajs])
    for lo, hi in ((1, 16), (17, 32), (33, 48), (49, 64)):
        sel = (ncs >= lo) & (ncs <= hi)
        log(f"  n_c {lo:2d}-{hi:2d}: busy mean {busy[sel].mean() * 1e3:.1f} ms,"
            f" max {busy[sel].max() * 1e3:.1f} ms")
    # load per SIMD: how many trajectories each ran, and its longest busy time
    keys, per_simd = np.unique(res["simd"], return_counts=True)
    for c in np.unique(per_simd):
        on = np.isin(res["simd"], keys[per_simd == c])
        log(f"  {int((per_simd == c).sum())} SIMDs ran {c} trajectories: "
            f"busy mean {busy[on].mean() * 1e3:.1f} ms, "
            f"max {busy[on].max() * 1e3:.1f} ms")
    resident = min(wpc, bt.resident_per_cu())
    from pycd_amd import _native
    cus = _native.device_cus()
    log(f"  kernel {bt_k:.1f} ms; longest trajectory {busy.max() * 1e3:.1f} ms; "
        f"busy sum / resident waves ({resident}/CU) "
        f"{busy.sum() / (resident * cus) * 1e3:.1f} ms")
    algo = algorithmic_bytes_batched(trajs, fams, hops, res["start"], 8)
    achieved = algo / (bt_k / 1e3) / 1e9
    info = {
        "ms_per_step": ms, "total_hops": total_hops, "hops_per_s": value,
        "waves_per_cu": wpc, "no_select": int(res["no_select"].sum()),
        "longest_trajectory_ms": float(busy.max() * 1e3),
        "resident_waves_per_cu": resident,
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.traffic_profile),
                     "kernel": "kmc_batched_kernel<double,true>",
                     "launch_ms": bt_k, "algorithmic_bytes": algo},
    }
    bt.close()
    return info, trajs, fams


def run_plain(args, log):
    from pycd_amd import workloads, _native, rng
    from pycd_amd.rng import seeded_starts
    w = workloads.corundum(8, n_carriers=16, field_vpm=1e6, t_final=2e-8)
    table = w.build_table()
    run = w.make_run(table)
    pb = _native.KmcProblem(
        table=table, lattice_q=run.base_charge_config()[:, 0],
        nbr_count=run.nbr_count, nbr_site=run.nbr_site, nbr_vec=run.nbr_vec,
        nbr_lambda=run.nbr_lambda, nbr_vab=run.nbr_vab,
        site_energy=run.system_relative_energies,
        carrier_q=np.asarray(run.species_charge_list, dtype=np.float64),
        field=run.electric_field, nu=run.material.vn, kT=run.temp,
        time_interval=run.time_interval,
        n_path=int(run.t_final / run.time_interval) + 1,
        cells=run.system_size)
    sess = _native.KmcSession(pb)
    if args.plain_variant != "auto":
        sess.set_variant(args.plain_variant)
    variant = sess.variant()
    log(f"plain: {variant} kernel; hop loop reads {sess.table_rows()} "
        f"table rows of {w.n_sites}")
    occs, words = seeded_starts(2, 1, [run.hop_site_lists["electron"]], [16])
    state = (3, tuple(int(v) for v in words[0]), None)
    n = args.plain_hops
    uni = rng.UniformStream(state).take(2 * n)
    dev = _native.DeviceArray.from_host(uni)
    sess.begin(occs[0])
    ms = sess.advance_device(dev, n)
    hops = sess.read()["hops"]
    sess.close()
    dev.free()
    nc = 16
    z = int(np.max(run.nbr_count))
    per_hop = plain_bytes_per_hop(variant, nc, z)
    gbs = per_hop * hops / (ms / 1e3) / 1e9
    _KEEP["plain"] = (run, table)
    return {"hops": hops, "ms": ms, "hops_per_s": hops / (ms / 1e3),
            "n_sites": w.n_sites, "n_carriers": nc, "z": z,
            "variant": variant,
            "roofline": {"bound": "hbm", "achieved": gbs,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS,
                         "kernel": ("kmc_wave_kernel" if variant == "wave"
                                    else "kmc_plain_kernel"),
                         "algorithmic_bytes_per_hop": per_hop}}


def plain_bytes_per_hop(variant, nc, z):
    """Bytes the plain kernel's algorithm reads per hop.

    wave (row cache in LDS): (1 + z)(2 n_c - 1) table entries (column m of
    the other rows, the mover's 1 + z new rows), 10 doubles of rate inputs
    for each of the mover's z new processes (phi_L and E at both ends, the
    self term, lambda, V_ab, hop vector), 2 z^2 words of look-ahead slot
    tables and 24 B of uniforms (u1, u2, log u2).
    workgroup: n_c (1 + z) x n_c table entries (every row gathered again),
    11 doubles per process and 16 B of uniforms."""
    if variant == "wave":
        return 8 * (1 + z) * (2 * nc - 1) + 80 * z + 8 * z * z + 24
    return 8 * nc * (1 + z) * nc + 8 * 11 * nc * z + 16


def run_serve(args, log):
    """Serving form: Poisson arrivals (200 jobs/s) of monoclinic-lattice jobs
    (edges U{4..12}: 256-6912 sites, 1-32 carriers, hops log-uniform
    1e5-1e6) through the continuous-batching server; tables of all nine
    sizes are built once before the clock starts (the cache then serves
    repeat sizes)."""
    from pycd_amd import workloads
    from pycd_amd.serve import KmcServer, poisson_jobs
    cache = {}

    def family(edge):
        if edge not in cache:
            w = workloads.monoclinic(edge)
            table = w.build_table()
            cache[edge] = (w.make_run(table), table)
        return cache[edge]

    t0 = time.time()
    srv = KmcServer(family, n_slots=2048, nc_max=32, z_max=2,
                    rows_capacity=2048 * 801 * 96, slice_hops=args.serve_slice,
                    rate_precision="f32", path_rows=200, msd_lags=100,
                    waves_per_cu=8)
    for edge in range(4, 13):
        srv._family_for(edge)
    cold = time.time() - t0
    log(f"serve: 9 table families built in {cold:.1f}s")
    warm = poisson_jobs(32, rate=1e6, hops=(2e3, 4e3), seed=1)
    srv.serve(warm)
    jobs = poisson_jobs(args.serve_jobs, rate=200.0, seed=5)
    srv.launch_ms.clear()
    stats = srv.serve(jobs)
    stats["cold_start_s"] = cold
    stats["arrival_rate_per_s"] = 200.0
    srv.close()
    return stats


def run_precision(args, log):
    """Precision form: 10x10x10 corundum (12000 sites), reciprocal sum with
    |k| < 2 pi 10 / L_a, table in FP64 / FP32-MFMA / BF16x3-MFMA; per-entry
    and total-energy error against FP64; then 1e6 hops with 32 electrons on
    each table (plain path), compared with the FP64-table run."""
    from pycd_amd import workloads, _native, rng, constants, tuning
    from pycd_amd.rng import seeded_starts
    n = args.precision_edge
    w = workloads.corundum(n, m_kcut=n, n_carriers=32, t_final=1e-8)
    _KEEP["precision"] = w
    pos = w.neighbors.bulk_sites.cell_coordinates
    out = {"n_sites": int(w.n_sites), "autotune": {}}
    tables = {}
    targs = (pos, w.neighbors.simulation_cell_matrix,
             w.system.translational_matrix, w.system.pbc, w.system.alpha,
             w.system.r_cut, w.system.k_cut, 1.0)
    for mode in ("f64", "f32", "bf16x3"):
        # the tile variant for this lattice, by the autotuner
        v, times = tuning.autotune_table(*targs, precision=mode, repeats=1,
                                         log=log)
        out["autotune"][mode] = {"variant": v,
                                 "ms": {str(k): t for k, t in times.items()}}
        best = None
        for rep_i in range(2):
            tab, rep = _native.pair_table(
                pos, w.neighbors.simulation_cell_matrix,
                w.system.translational_matrix, w.system.pbc, w.system.alpha,
                w.system.r_cut, w.system.k_cut, 1.0, mode)
            if best is None or rep["ms_recip"] < best["ms_recip"]:
                best = rep
        tables[mode] = tab
        # the symmetric product: N (N + 1) / 2 entries of 2 Kp flops each
        kp = 2 * best["n_kvec"]
        nn = w.n_sites
        flops = float(nn) * (nn + 1) * kp * (3 if mode == "bf16x3" else 1)
        peak = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0}[mode]
        tf = flops / (best["ms_recip"] / 1e3) / 1e12
        out[mode] = {"ms_sf": best["ms_sf"], "ms_contraction": best["ms_recip"],
                     "n_kvec_half": best["n_kvec"], "k_max": best["k_max"],
                     "tflops": tf, "peak_tflops": peak, "frac": tf / peak}
        if best["ms_real"] > 0:
            # the real-space term in a kernel of its own (inside
            # ms_contraction): the contraction kernel's own rate beside it
            ms_k = best["ms_recip"] - best["ms_real"]
            out[mode]["ms_real_kernel"] = best["ms_real"]
            out[mode]["frac_contraction_kernel"] = (
                flops / (ms_k / 1e3) / 1e12 / peak)
        log(f"{mode}: {best['ms_recip']:.2f} ms contraction, {tf:.1f} TF/s")
    ref = tables["f64"]
    scale = float(np.max(np.abs(ref)))
    run = w.make_run(ref)
    occs, words = seeded_starts(2, 1, [run.hop_site_lists["electron"]], [32])
    q = run.base_charge_config()[:, 0].copy()
    q[occs[0]] -= 1.0
    e_ref = float(q @ ref @ q)
    for mode in ("f32", "bf16x3"):
        d = tables[mode] - ref
        out[mode]["max_abs_err"] = float(np.max(np.abs(d)))
        out[mode]["rms_err"] = float(np.sqrt(np.mean(d * d)))
        out[mode]["rel_max_err"] = out[mode]["max_abs_err"] / scale
        out[mode]["energy_err_hartree"] = float(abs(q @ d @ q))
        out[mode]["energy_rel_err"] = out[mode]["energy_err_hartree"] / abs(e_ref)
        del d
    out["energy_f64_hartree"] = e_ref
    state = (3, tuple(int(v) for v in words[0]), None)
    hops = args.precision_hops
    uni = rng.UniformStream(state).take(2 * hops)
    dev = _native.DeviceArray.from_host(uni)
    runs = {}
    for mode in ("f64", "f32", "bf16x3"):
        pb = _native.KmcProblem(
            table=tables[mode], lattice_q=run.base_charge_config()[:, 0],
            nbr_count=run.nbr_count, nbr_site=run.nbr_site,
            nbr_vec=run.nbr_vec, nbr_lambda=run.nbr_lambda,
            nbr_vab=run.nbr_vab, site_energy=run.system_relative_energies,
            carrier_q=np.full(32, -1.0), field=np.zeros(3),
            nu=run.material.vn, kT=run.temp, time_interval=run.time_interval,
            n_path=int(run.t_final / run.time_interval) + 1)
        sess = _native.KmcSession(pb)
        sess.begin(occs[0])
        ms = sess.advance_device(dev, hops, max_hops=hops)
        r = sess.read()
        sess.close()
        runs[mode] = r
        out[mode]["kmc_hops"] = r["hops"]
        out[mode]["kmc_hops_per_s"] = r["hops"] / (ms / 1e3)
        out[mode]["kmc_final_time_s"] = r["t"] / constants.SEC2AUTIME
        log(f"{mode}: {r['hops']} hops, {r['hops'] / (ms / 1e3):.3e} hops/s")
    for mode in ("f32", "bf16x3"):
        a, b = runs[mode]["unwrapped"], runs["f64"]["unwrapped"]
        rows = min(runs[mode]["start"], runs["f64"]["start"])
        same = np.all(np.abs(a[:rows] - b[:rows]) < 1e-6, axis=1)
        first = int(np.argmin(same)) if not same.all() else rows
        out[mode]["path_rows_identical_to_f64"] = first
        out[mode]["final_time_rel_diff"] = float(
            abs(runs[mode]["t"] - runs["f64"]["t"]) / runs["f64"]["t"])
    dev.free()
    return out


def _oracle_rate(pb, occ, gen, budget):
    """Steady-state seconds per hop of the NumPy oracle on one trajectory,
    run for about `budget` seconds: the per-trajectory set-up (initial energy
    q^T A q) is timed with max_hops=0 and subtracted.  Returns (hops, s)."""
    from oracle import kmc
    t0 = time.perf_counter()
    kmc.run(pb, occ, gen.random, max_hops=0)
    setup = time.perf_counter() - t0
    # time per hop from the difference of two short runs (the set-up, in
    # both, cancels), so that noise in the set-up time cannot shrink it
    t0 = time.perf_counter()
    kmc.run(pb, occ, gen.random, max_hops=50)
    t1 = time.perf_counter()
    kmc.run(pb, occ, gen.random, max_hops=250)
    t2 = time.perf_counter()
    per = max(((t2 - t1) - (t1 - t0)) / 200, 1e-7)
    n = max(50, int(budget / per))
    t0 = time.perf_counter()
    res = kmc.run(pb, occ, gen.random, max_hops=n)
    return res["hops"], max(time.perf_counter() - t0 - setup, 1e-9)


def _oracle_inputs(t, fam, k):
    import random
    from oracle import kmc
    run = fam.run
    pb = kmc.Problem(fam.table, fam.lattice_q, run.nbr_count, run.nbr_site,
                     run.nbr_vec, run.nbr_lambda, run.nbr_vab,
                     run.system_relative_energies, [-1.0] * t["n_carriers"],
                     t["field"], fam.nu, t["kT"], t["time_interval"],
                     t["n_path"])
    gen = random.Random()
    gen.seed(int(k))
    occ = gen.sample(fam.hop_sites, t["n_carriers"])
    return pb, occ, gen


def cpu_worker(spec, budget):
    """One worker of the all-cores baseline: the oracle on the trajectory
    described by the .npz `spec` (its table memory-mapped from the .npy it
    names), for about `budget` seconds; prints hops and seconds as JSON."""
    import random
    from oracle import kmc
    z = np.load(spec, allow_pickle=False)
    table = np.load(str(z["table"]), mmap_mode="r")
    pb = kmc.Problem(table, z["lattice_q"], z["count"], z["nbr"], z["vec"],
                     z["lam"], z["vab"], z["energy"], z["carrier_q"],
                     z["field"], float(z["nu"]), float(z["kT"]),
                     float(z["dt"]), int(z["n_path"]))
    gen = random.Random()
    gen.seed(int(z["seed"]))
    occ = gen.sample([int(v) for v in z["hop_sites"]], int(z["nc"]))
    hops, el = _oracle_rate(pb, occ, gen, float(budget))
    print(json.dumps({"hops": int(hops), "seconds": el}), flush=True)


def _oracle_spec(path, table_path, run, lattice_q, nc, field, nu, kT, dt,
                 n_path, seed, hop_sites):
    np.savez(path, table=table_path, lattice_q=lattice_q,
             count=run.nbr_count, nbr=run.nbr_site, vec=run.nbr_vec,
             lam=run.nbr_lambda, vab=run.nbr_vab,
             energy=run.system_relative_energies,
             carrier_q=np.full(nc, -1.0), field=np.asarray(field), nu=nu,
             kT=kT, dt=dt, n_path=n_path, seed=int(seed),
             hop_sites=np.asarray(hop_sites), nc=nc)


def _run_workers(args_list, entry, budget, log, what):
    """One worker process (one thread) per spec, all at once; each prints
    {"hops"/"rows": n, "seconds": s} as JSON.  Returns the parsed lines."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1")
    procs = [subprocess.Popen(
        [sys.executable, "-c", f"import bench; bench.{entry}({a})"],
        cwd=str(Path(__file__).resolve().parent), env=env,
        stdout=subprocess.PIPE, text=True) for a in args_list]
    log(f"cpu {what}: {len(procs)} workers started")
    out = []
    for i, p in enumerate(procs):
        try:
            so, _ = p.communicate(timeout=4 * budget + 60)
        except subprocess.TimeoutExpired:
            p.kill()
            p.communicate()
            log(f"cpu {what} worker {i} timed out; left out")
            continue
        if p.returncode != 0:
            raise RuntimeError(f"cpu worker exited with {p.returncode}")
        out.append(json.loads(so.strip().splitlines()[-1]))
    return out


def cpu_all_cores(args, trajs, fams, cores, log):
    """The oracle on `cores` sampled trajectories at once, one worker
    process (one thread) per core.  Every trajectory of the workload makes
    the same number of hops, so cores working through it run at
    cores / mean(time per hop) over the sample."""
    import tempfile
    rng_ = np.random.default_rng(98)
    sample = rng_.choice(len(trajs), size=min(cores, len(trajs)),
                         replace=False)
    # tables memory-mapped by the workers: kept in memory (/dev/shm) when
    # the host has it
    shm = "/dev/shm" if os.path.isdir("/dev/shm") else None
    with tempfile.TemporaryDirectory(dir=shm) as tmp:
        tables, specs = {}, []
        for k in sample:
            t = trajs[k]
            f = t["family"]
            fam = fams[f]
            if f not in tables:
                tables[f] = os.path.join(tmp, f"table{f}.npy")
                np.save(tables[f], np.ascontiguousarray(fam.table))
            spec = os.path.join(tmp, f"traj{k}.npz")
            _oracle_spec(spec, tables[f], fam.run, fam.lattice_q,
                         t["n_carriers"], t["field"], fam.nu, t["kT"],
                         t["time_interval"], t["n_path"], k, fam.hop_sites)
            specs.append(f"{spec!r}, {args.cpu_seconds!r}")
        res = _run_workers(specs, "cpu_worker", args.cpu_seconds, log,
                           "batched")
    per_hop = [r["seconds"] / max(r["hops"], 1) for r in res]
    hops = sum(r["hops"] for r in res)
    total = len(per_hop) / float(np.mean(per_hop))
    log(f"cpu all cores: {len(per_hop)} workers, {total:.1f} hops/s")
    return total, hops, len(per_hop)


def cpu_plain(args, cores, log):
    """The oracle on the plain configuration (8x8x8, 6144 sites, 16
    electrons, 1e6 V/m): one trajectory on one core, then `cores`
    trajectories (seeds 2, 3, ...) at once, one per core."""
    import random
    import tempfile
    from threadpoolctl import threadpool_limits
    from oracle import kmc, seeding
    run, table = _KEEP["plain"]
    budget = args.cpu_seconds / 2
    pb = kmc.Problem(table, run.base_charge_config()[:, 0], run.nbr_count,
                     run.nbr_site, run.nbr_vec, run.nbr_lambda, run.nbr_vab,
                     run.system_relative_energies, [-1.0] * 16,
                     run.electric_field, run.material.vn, run.temp,
                     run.time_interval,
                     int(run.t_final / run.time_interval) + 1)
    gen = random.Random()
    gen.setstate(seeding.trajectory_states(2, 1)[0])
    occ = seeding.initial_occupancy(gen, run.hop_site_lists["electron"], 16)
    with threadpool_limits(1):
        hops, el = _oracle_rate(pb, occ, gen, budget)
    out = {"one_core": hops / el, "one_core_sample": f"{hops} hops in "
                                                     f"{el:.1f} s"}
    log(f"cpu plain one core: {hops} hops in {el:.1f} s")
    if cores > 1 and not args.cpu_single:
        shm = "/dev/shm" if os.path.isdir("/dev/shm") else None
        with tempfile.TemporaryDirectory(dir=shm) as tmp:
            tpath = os.path.join(tmp, "table.npy")
            np.save(tpath, np.ascontiguousarray(table))
            specs = []
            for c in range(cores):
                spec = os.path.join(tmp, f"plain{c}.npz")
                _oracle_spec(spec, tpath, run, run.base_charge_config()[:, 0],
                             16, run.electric_field, run.material.vn,
                             run.temp, run.time_interval,
                             int(run.t_final / run.time_interval) + 1, 2 + c,
                             run.hop_site_lists["electron"])
                specs.append(f"{spec!r}, {budget!r}")
            res = _run_workers(specs, "cpu_worker", budget, log, "plain")
        per_hop = [r["seconds"] / max(r["hops"], 1) for r in res]
        out["all_cores"] = len(per_hop) / float(np.mean(per_hop))
        out["cores"] = len(per_hop)
        log(f"cpu plain all cores: {out['all_cores']:.1f} hops/s")
    return out


def table_worker(spec, n_rows):
    """Rows of the oracle's FP64 pair table at the precision configuration
    from the .npz `spec`; prints rows and seconds as JSON."""
    from oracle import ewald
    z = np.load(spec, allow_pickle=False)
    rows = [int(r) for r in z["rows"][:n_rows]]
    t0 = time.perf_counter()
    ewald.table_rows(z["pos"], rows, z["lattice"], z["size"], z["pbc"],
                     float(z["alpha"]), float(z["r_cut"]), float(z["k_cut"]),
                     float(z["eps"]))
    print(json.dumps({"rows": len(rows),
                      "seconds": time.perf_counter() - t0}), flush=True)


def cpu_table(args, cores, log):
    """Build time of the precision configuration's FP64 table (12000 sites,
    K_half = 4918) by the oracle, from sampled rows: seconds per row on one
    core, and rows per second with one worker per core, each scaled to the
    full N rows."""
    import tempfile
    w = _KEEP["precision"]
    s = w.system
    n = int(w.n_sites)
    shm = "/dev/shm" if os.path.isdir("/dev/shm") else None
    out = {"n_sites": n}
    with tempfile.TemporaryDirectory(dir=shm) as tmp:
        rng_ = np.random.default_rng(7)
        specs = []
        for c in range(max(cores, 1)):
            spec = os.path.join(tmp, f"rows{c}.npz")
            np.savez(spec, pos=w.neighbors.bulk_sites.cell_coordinates,
                     rows=rng_.choice(n, size=64, replace=False),
                     lattice=w.material.lattice_matrix, size=w.size,
                     pbc=np.asarray(w.neighbors.pbc), alpha=s.alpha,
                     r_cut=s.r_cut, k_cut=s.k_cut,
                     eps=w.material.dielectric_constant)
            specs.append(spec)
        # one core: a few rows, timed in a worker of its own (one thread)
        one = _run_workers([f"{specs[0]!r}, 6"], "table_worker", 30, log,
                           "table one core")[0]
        per_row = one["seconds"] / one["rows"]
        out["one_core_s_per_row"] = per_row
        out["one_core_full_table_s"] = per_row * n
        log(f"cpu table one core: {per_row:.3f} s per row, "
            f"{per_row * n:.0f} s for {n} rows")
        if cores > 1 and not args.cpu_single:
            k = max(2, int(min(64, args.cpu_seconds / 2 / per_row)))
            res = _run_workers([f"{sp!r}, {k}" for sp in specs],
                               "table_worker", k * per_row * 4 + 30, log,
                               "table all cores")
            rate = sum(r["rows"] / r["seconds"] for r in res)
            out["all_cores_rows_per_s"] = rate
            out["all_cores_full_table_s"] = n / rate
            out["cores"] = len(res)
            log(f"cpu table all cores: {n / rate:.0f} s for {n} rows")
    return out


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(args, trajs, fams, log):
    """NumPy oracle on a bounded sample of the batched workload: one core,
    then one trajectory per host core at once; beside it the plain
    configuration and the precision configuration's table build."""
    from threadpoolctl import threadpool_limits
    rng_ = np.random.default_rng(99)
    sample = rng_.choice(len(trajs), size=min(8, len(trajs)), replace=False)
    budget = args.cpu_seconds / len(sample)
    per_hop, done_hops, spent = [], 0, 0.0
    for k in sample:
        t = trajs[k]
        pb, occ, gen = _oracle_inputs(t, fams[t["family"]], k)
        with threadpool_limits(1):
            hops, el = _oracle_rate(pb, occ, gen, budget)
        spent += el
        done_hops += hops
        per_hop.append(el / max(hops, 1))
        log(f"cpu one core: trajectory {int(k)}, {hops} hops in {el:.1f} s")
    # every trajectory of the workload makes the same number of hops, so the
    # workload rate is 1 / mean(time per hop) over the sampled trajectories
    one = 1.0 / float(np.mean(per_hop))
    out = {"value": one, "unit": "hops/s", "cores": 1, "kind": "port",
           "single_core": one,
           "sample": f"{len(sample)} trajectories of the batched workload "
                     f"(equal hop counts), {done_hops} hops in {spent:.1f} "
                     f"s, NumPy oracle of Run.do_kmc_steps, 1 thread; "
                     f"rate = 1 / mean time per hop"}
    # the host's cores (the box's share: at most 16), one trajectory each
    cores = min(len(os.sched_getaffinity(0)), 16)
    if cores > 1 and not args.cpu_single:
        total, hops, n = cpu_all_cores(args, trajs, fams, cores, log)
        out.update(value=total, cores=n, sample=(
            f"{n} trajectories of the batched workload at once, one worker "
            f"process per core, about {args.cpu_seconds:.0f} s each "
            f"({hops} hops), NumPy oracle of Run.do_kmc_steps, 1 thread per "
            f"worker; value = cores / mean steady-state time per hop; "
            f"single_core = "
            f"{len(sample)} trajectories one after another on one core "
            f"({done_hops} hops in {spent:.1f} s)"))
    out["cpu_model"] = cpu_model()
    out["host_cores"] = len(os.sched_getaffinity(0))
    if "plain" in _KEEP:
        out["plain_cfg2"] = cpu_plain(args, cores, log)
    if "precision" in _KEEP:
        out["table_cfg3"] = cpu_table(args, cores, log)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              f"trajectories are independent and would be sharded across "
              f"GPUs by trajectory",
              file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.copy_calibration > 0:
        from pycd_amd import _native
        n = args.copy_calibration // 8
        a = _native.DeviceArray(n * 8, np.float64, (n,))
        b = _native.DeviceArray(n * 8, np.float64, (n,))
        for _ in range(3):
            _native.lib().pycd_memcpy(a.ctx, b.ptr, a.ptr, n * 8)
        a.free()
        b.free()
        log(f"calibration: 3 device copies of {n * 8} bytes")
    if args.mode == "precision":
        print(json.dumps({"mode": "precision",
                          "result": run_precision(args, log)},
                         default=_json_default))
        return
    if args.mode == "plain":
        print(json.dumps({"mode": "plain", "result": run_plain(args, log)},
                         default=_json_default))
        return
    info, trajs, fams = run_batched(args, log)
    log(f"batched: {info['hops_per_s']:.4e} hops/s, {info['ms_per_step']:.1f}"
        f" ms/step")
    plain = None
    if not args.no_plain:
        plain = run_plain(args, log)
        log(f"plain: {plain['hops_per_s']:.4e} hops/s")
    serve = None
    if not args.no_serve:
        serve = run_serve(args, log)
        log(f"serve: {serve['hops_per_s']:.4e} hops/s, p50 "
            f"{serve['p50_latency_ms']:.0f} ms, p99 "
            f"{serve['p99_latency_ms']:.0f} ms")
    precision = None
    if not args.no_precision:
        precision = run_precision(args, log)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, trajs, fams, log)
        log(f"cpu: {cpu['value']:.1f} hops/s")
    line = {
        "metric": "KMC hops/s summed, batched vs plain on 1 GPU; p50/p99 "
                  "job latency",
        "value": info["hops_per_s"],
        "unit": "hops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": info["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 rates, f64 table/prefix/clock",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {len(trajs)} trajectories, corundum "
                        f"edge 4-8 (768-6144 sites), 1-64 electrons, "
                        f"|E| 0-1e7 V/m, T 250-400 K, {args.hops} hops each",
            "global_batch": len(trajs), "seq_len": args.hops,
            "parallelism": "1 GPU, one wavefront per trajectory",
            "waves_per_cu": info["waves_per_cu"],
        },
        "plain_hops_per_s": plain["hops_per_s"] if plain else None,
        "plain_roofline": plain["roofline"] if plain else None,
        "batched_over_plain": (info["hops_per_s"] / plain["hops_per_s"]
                               if plain else None),
        "p50_latency_ms": serve["p50_latency_ms"] if serve else None,
        "p99_latency_ms": serve["p99_latency_ms"] if serve else None,
        "serve": serve,
        "total_hops": info["total_hops"],
        "longest_trajectory_ms": info["longest_trajectory_ms"],
        "resident_waves_per_cu": info["resident_waves_per_cu"],
        "roofline": info["roofline"],
        "precision": precision,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line, default=_json_default))


if __name__ == "__main__":
    main()

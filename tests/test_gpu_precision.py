"""Reduced-precision reciprocal sum (FP32 MFMA, BF16x3 MFMA) vs FP64.

Tolerances (written here, measured numbers in the printout):
  per entry   max |A_mode - A_f64| / max |A_f64|  <= 1e-5 (f32), 1e-4 (bf16x3)
  energy      |q^T (A_mode - A_f64) q| / |q^T A_f64 q| <= 1e-5 / 1e-4
"""
import numpy as np
import pytest

from pycd_amd import _native, workloads

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-5, "bf16x3": 1e-4}


def _tables(w, pos):
    out = {}
    for mode in ("f64", "f32", "bf16x3"):
        out[mode], rep = _native.pair_table(
            pos, w.neighbors.simulation_cell_matrix,
            w.system.translational_matrix, w.system.pbc, w.system.alpha,
            w.system.r_cut, w.system.k_cut, 1.0, mode)
        assert rep["precision"] == mode
    return out


@pytest.mark.parametrize("size", [3, 6])
def test_modes_against_f64(size):
    w = workloads.corundum(size)
    pos = w.neighbors.bulk_sites.cell_coordinates
    tabs = _tables(w, pos)
    ref = tabs["f64"]
    scale = np.max(np.abs(ref))
    run = w.make_run(ref)
    q = run.base_charge_config()[:, 0].copy()
    rng = np.random.default_rng(3)
    occ = rng.choice(run.hop_site_lists["electron"], size=8, replace=False)
    q[occ] -= 1.0
    e_ref = q @ ref @ q
    for mode in ("f32", "bf16x3"):
        d = tabs[mode] - ref
        err_max = np.max(np.abs(d)) / scale
        err_rms = np.sqrt(np.mean(d * d)) / scale
        e_err = abs(q @ d @ q) / abs(e_ref)
        print(f"N={len(pos)} {mode}: max {err_max:.3e} rms {err_rms:.3e} "
              f"energy {e_err:.3e}")
        assert err_max <= TOL[mode]
        assert e_err <= TOL[mode]
        # the reduced precision touches only the reciprocal part: symmetric
        assert np.max(np.abs(tabs[mode] - tabs[mode].T)) <= 1e-6 * scale


# bounds at the precision configuration (12000 sites); measured there:
# max entry error 1.4e-6 (f32) and 2.0e-6 (bf16x3) of max|A|, total energy
# 6.6e-7 and 7.5e-5 relative.  BF16x3 keeps about 16 bits of each operand
# (8 in hi, 8 in lo), FP32 24.
ENTRY_TOL_12000 = {"f32": 1e-5, "bf16x3": 1e-5}
ENERGY_TOL_12000 = {"f32": 5e-6, "bf16x3": 5e-4}


def test_precision_configuration_10x10x10():
    """10x10x10 corundum (12000 sites), |k| < 2 pi 10 / L_a: K_half = 4918
    and k_max (9, 9, 28) as the reference's half space gives them; eight
    full FP64 rows against the oracle at 1e-12 of the table's scale; the
    FP32 and BF16x3 tables against FP64 within the bounds above, the BF16x3
    one equal bit for bit across its kernels; then
    20000 hops of 32 electrons on each table, with the energy bookkeeping
    equal to q^T A q of the final state on that table."""
    from oracle import ewald
    from tests.kmc_helpers import native_problem, run_gpu
    w = workloads.corundum(10, m_kcut=10, n_carriers=32, t_final=1e-8)
    pos = w.neighbors.bulk_sites.cell_coordinates
    assert len(pos) == 12000
    tabs, reps = {}, {}
    for mode in ("f64", "f32", "bf16x3"):
        tabs[mode], reps[mode] = _native.pair_table(
            pos, w.neighbors.simulation_cell_matrix,
            w.system.translational_matrix, w.system.pbc, w.system.alpha,
            w.system.r_cut, w.system.k_cut, 1.0, mode)
        assert reps[mode]["n_kvec"] == 4918
        assert list(reps[mode]["k_max"]) == [9, 9, 28]
    # the default BF16x3 kernel here: 1024 256 x 256 tiles (four whole
    # rounds of the CUs) and 104 more as 416 128 x 128 tiles in pairs; the
    # same table bit for bit as the 128 x 128 two-stage kernel
    try:
        _native.set_table_variant("bf16x3", 3)
        alt, _ = _native.pair_table(
            pos, w.neighbors.simulation_cell_matrix,
            w.system.translational_matrix, w.system.pbc, w.system.alpha,
            w.system.r_cut, w.system.k_cut, 1.0, "bf16x3")
    finally:
        _native.set_table_variant("bf16x3", 0)
    assert np.array_equal(alt, tabs["bf16x3"])
    del alt
    ref = tabs["f64"]
    scale = np.max(np.abs(ref))
    rows = [0, 1, 11, 4321, 5999, 6000, 11988, 11999]
    s = w.system
    ora = ewald.table_rows(pos, rows, w.material.lattice_matrix, w.size,
                           w.neighbors.pbc, s.alpha, s.r_cut, s.k_cut, 1.0)
    dev = np.max(np.abs(ref[rows] - ora)) / scale
    print(f"f64 rows vs oracle: {dev:.2e} of max|A|")
    assert dev <= 1e-12
    run = w.make_run(ref)
    rng = np.random.default_rng(4)
    occ = rng.choice(run.hop_site_lists["electron"], size=32, replace=False)
    q = run.base_charge_config()[:, 0].copy()
    q[occ] -= 1.0
    e_ref = q @ ref @ q
    for mode in ("f32", "bf16x3"):
        d = tabs[mode] - ref
        err_max = np.max(np.abs(d)) / scale
        err_rms = np.sqrt(np.mean(d * d)) / scale
        e_err = abs(q @ d @ q) / abs(e_ref)
        print(f"{mode}: max {err_max:.3e} rms {err_rms:.3e} energy "
              f"{e_err:.3e} ({reps[mode]['ms_recip']:.1f} ms)")
        assert err_max <= ENTRY_TOL_12000[mode]
        assert e_err <= ENERGY_TOL_12000[mode]
        del d
    n = 20000
    clocks = {}
    for mode in ("f64", "f32", "bf16x3"):
        table = tabs[mode]
        res, occ0 = run_gpu(run, table, 2, 32, max_hops=n, record=False)
        assert res["hops"] == n
        assert np.all(np.diff(res["time_data"]) > 0)
        qa = run.base_charge_config()[:, 0].copy()
        np.add.at(qa, np.asarray(occ0), -1.0)
        qb = run.base_charge_config()[:, 0].copy()
        np.add.at(qb, np.asarray(res["occupancy"]), -1.0)
        book = res["energy_now"] - res["energy"][0]
        direct = qb @ table @ qb - qa @ table @ qa
        # The reference's dG0 (PyCD/core.py:2004-2017) is the exact energy
        # change E(after) - E(before) when the table is symmetric with one
        # diagonal value; otherwise each hop s -> n is off by
        # q^2 (A[n, n] - A[s, s]) plus the asymmetry of the two rows.  The
        # FP64 table has both to rounding; the reduced-precision ones to
        # their own error, which bounds the bookkeeping over n hops.
        if mode == "f64":
            bound = 1e-8 * max(1.0, abs(direct))
        else:
            diag = np.diag(table)
            rows = [int(v) for v in np.unique(
                np.concatenate([occ0, res["occupancy"]]))]
            asym = np.max(np.abs(table[rows] - table[:, rows].T))
            bound = n * (np.ptp(diag) + 4 * 32 * asym)
        print(f"{mode}: bookkeeping {abs(book - direct):.3e} (bound "
              f"{bound:.3e})")
        assert abs(book - direct) < bound, mode
        clocks[mode] = res["time_data"][-1]
    for mode in ("f32", "bf16x3"):
        rel = abs(clocks[mode] / clocks["f64"] - 1)
        print(f"{mode}: {n} hops, clock {rel:.2e} from the FP64 table's")
        assert rel < 0.05


def test_bf16x3_tile_pairs_with_an_odd_count():
    """Monoclinic 12 x 12 x 12 (6912 sites): on 256 CUs, 378 256 x 256
    tiles, 256 of them whole and the other 122 as 479 128 x 128 tiles in
    pairs, the last pair's second half repeating the first and storing
    nothing.  The BF16x3 table equals the 128 x 128 two-stage kernel's bit
    for bit."""
    w = workloads.monoclinic(12)
    pos = w.neighbors.bulk_sites.cell_coordinates
    assert len(pos) == 6912
    args = (pos, w.neighbors.simulation_cell_matrix,
            w.system.translational_matrix, w.system.pbc, w.system.alpha,
            w.system.r_cut, w.system.k_cut, 1.0)
    tabs = []
    try:
        for v in (4, 3):
            _native.set_table_variant("bf16x3", v)
            tabs.append(_native.pair_table(*args, "bf16x3")[0])
    finally:
        _native.set_table_variant("bf16x3", 0)
    assert np.array_equal(tabs[0], tabs[1])
    assert np.all(np.isfinite(tabs[0]))


def test_table_variants_agree_and_autotune_picks_one():
    """Each precision's tile variants (FP64: 64 x 64 and 128 x 128 tiles;
    BF16x3: K slices of 32 and 64, and of 32 in two LDS stages; 2592 sites,
    not a whole number of 128-row tiles) give the same table bit for bit,
    and the autotuner keeps the fastest of them in the context."""
    from pycd_amd import tuning
    w = workloads.corundum(6)
    pos = w.neighbors.bulk_sites.cell_coordinates
    args = (pos, w.neighbors.simulation_cell_matrix,
            w.system.translational_matrix, w.system.pbc, w.system.alpha,
            w.system.r_cut, w.system.k_cut, 1.0)
    try:
        for mode in ("f64", "bf16x3"):
            tabs = []
            for v in _native.TABLE_VARIANTS[mode]:
                _native.set_table_variant(mode, v)
                assert _native.table_variant(mode) == v
                tabs.append(_native.pair_table(*args, mode)[0].copy())
            for t in tabs[1:]:
                assert np.array_equal(t, tabs[0]), mode
            best, times = tuning.autotune_table(*args, precision=mode)
            print(mode, {k: round(v, 3) for k, v in times.items()}, best)
            assert best in _native.TABLE_VARIANTS[mode]
            assert _native.table_variant(mode) == best
            assert times[best] == min(times.values())
    finally:
        for mode in ("f64", "f32", "bf16x3"):
            _native.set_table_variant(mode, 0)

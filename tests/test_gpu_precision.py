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

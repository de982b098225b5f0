"""The exact-diffusivity oracle against closed forms."""
import numpy as np
import pytest

from oracle.diffusion import exact_diffusivity, over_kbt
from oracle.neighbors import hop_neighbor_list
from pycd_amd import workloads


def _run_tables(w):
    return w.make_run(np.zeros((w.n_sites, w.n_sites)))


def test_one_sublattice_reduces_to_naive_formula():
    """Simple cubic lattice, 6 equivalent neighbours: D = sum k d^2 / 6."""
    nbr = np.array([[1, 1, 1, 1, 1, 1], [0, 0, 0, 0, 0, 0]])
    a = 3.0
    d = np.array([[a, 0, 0], [-a, 0, 0], [0, a, 0], [0, -a, 0], [0, 0, a],
                  [0, 0, -a]], dtype=float)
    vec = np.stack([d, d])
    lam = np.full((2, 6), 0.03)
    vab = np.full((2, 6), 0.002)
    count = np.array([6, 6])
    sub = np.array([0, 0])
    nu, kT = 1e-3, 0.001
    D, _ = exact_diffusivity(nbr, vec, lam, vab, count, sub, nu, kT)
    k = nu * np.exp(-(0.03 / 4 - 0.002) / kT)
    # sum over the 6 hops of k d d^T / 2 = k a^2 I
    assert np.allclose(D, np.eye(3) * k * a * a, rtol=1e-12)


def test_corundum_value():
    """Fe corundum, lambda = 1 eV, V_ab = 0.2/0.1 eV, nu = 1.85e13 /s,
    300 K: independent evaluation gives D = 1.136162e-3 cm^2/s and
    D/kBT = 4.394872e-2 cm^2/Vs (rates 2.674289e12 and 5.588327e10 /s)."""
    w = workloads.corundum(3, hop_lists=hop_neighbor_list)
    run = _run_tables(w)
    sub = np.arange(w.n_sites) % 12
    D, dcm = exact_diffusivity(run.nbr_site, run.nbr_vec, run.nbr_lambda,
                               run.nbr_vab, run.nbr_count, sub,
                               run.material.vn, run.temp)
    assert abs(dcm / 1.136162e-3 - 1) < 2e-6
    assert abs(over_kbt(dcm, 300.0) / 4.394872e-2 - 1) < 2e-6


def test_drift_velocity_one_sublattice_closed_form():
    """Simple cubic lattice, field along x: v_x = a (k+ - k-) with
    k+- = nu exp(-(lambda/4 - V_ab -+ 0.5 E a)/kT), and no drift across the
    field; at zero field no drift at all."""
    from oracle.diffusion import exact_drift_velocity
    nbr = np.array([[1, 1, 1, 1, 1, 1], [0, 0, 0, 0, 0, 0]])
    a = 3.0
    d = np.array([[a, 0, 0], [-a, 0, 0], [0, a, 0], [0, -a, 0], [0, 0, a],
                  [0, 0, -a]], dtype=float)
    vec = np.stack([d, d])
    lam = np.full((2, 6), 0.03)
    vab = np.full((2, 6), 0.002)
    count = np.array([6, 6])
    sub = np.array([0, 0])
    nu, kT, E = 1e-3, 0.001, 2e-4
    v = exact_drift_velocity(nbr, vec, lam, vab, count, sub, nu, kT,
                             [E, 0.0, 0.0])
    kp = nu * np.exp(-((0.03 / 4 - 0.002) - 0.5 * E * a) / kT)
    km = nu * np.exp(-((0.03 / 4 - 0.002) + 0.5 * E * a) / kT)
    assert v[0] == pytest.approx(a * (kp - km), rel=1e-12)
    assert abs(v[1]) < 1e-12 * abs(v[0]) and abs(v[2]) < 1e-12 * abs(v[0])
    v0 = exact_drift_velocity(nbr, vec, lam, vab, count, sub, nu, kT,
                              [0.0, 0.0, 0.0])
    assert np.all(np.abs(v0) < 1e-15)

"""Oracle: `Analysis.compute_msd` numerics (test infrastructure only).

Restates PyCD/core.py:2996-3007 (per-lag windowed squared displacement,
`np.mean(np.einsum('ijk,ijk->ij', d, d), axis=0)`), the species and
trajectory means with SEM = std / sqrt(n_traj) (:3008-3036) and the D/kBT
conversion of the least-squares slope (:3055-3071).
"""
import numpy as np

from pycd_amd import constants


def squared_displacement(position_array, n_traj, n_path, n_lags):
    """position_array: (n_traj * P, n_c, 3) -> sd (n_traj, M, n_c)."""
    nc = position_array.shape[1]
    sd = np.zeros((n_traj, n_lags, nc))
    for t in range(n_traj):
        head = t * n_path
        for tau in range(1, n_lags):
            m = n_path - tau
            idx = np.arange(m)
            diff = position_array[head + tau + idx] - position_array[head + idx]
            sd[t, tau, :] = np.mean(np.einsum("ijk,ijk->ij", diff, diff),
                                    axis=0)
    return sd


def brute_force_sd(track, n_lags):
    """Double loop over origins for one (P, 3) track; tiny inputs only."""
    P = len(track)
    out = np.zeros(n_lags)
    for tau in range(1, n_lags):
        acc = 0.0
        for t0 in range(P - tau):
            d = track[t0 + tau] - track[t0]
            acc += float(d @ d)
        out[tau] = acc / (P - tau)
    return out


def diffusivity_over_kbt(slope_ang2_per_ns, n_dim, temp):
    """slope (A^2/ns) -> D/kBT in cm^2/(V s) (PyCD/core.py:3062-3065)."""
    kbt = constants.KB * temp / constants.EV2J
    return (slope_ang2_per_ns * constants.ANG2CM ** 2 * constants.SEC2NS
            / (2 * n_dim) / kbt)

"""HIP MSD reduction vs the oracle."""
import numpy as np
import pytest

from oracle import msd as omsd
from pycd_amd import _native, constants

pytestmark = pytest.mark.gpu


def test_shipped_trajectory(hematite):
    pos = hematite.unwrapped.reshape(-1, 1, 3) / constants.ANG2BOHR
    P = len(pos)
    M = 501
    ref = omsd.squared_displacement(pos, 1, P, M)
    got = _native.msd(pos.reshape(1, P, 1, 3), M)
    err = np.max(np.abs(got - ref) / np.maximum(ref, 1e-300))
    print("msd rel err", err)
    assert err < 1e-12


def test_random_multi_traj_multi_carrier():
    rng = np.random.default_rng(5)
    n_traj, P, nc, M = 3, 777, 5, 101
    pos = np.cumsum(rng.normal(size=(n_traj * P, nc, 3)), axis=0)
    ref = omsd.squared_displacement(pos, n_traj, P, M)
    got = _native.msd(pos.reshape(n_traj, P, nc, 3), M)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
    assert np.all(got[:, 0, :] == 0)

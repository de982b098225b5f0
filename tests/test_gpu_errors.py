"""Argument errors are refused on the host, with a message, before any
kernel runs: the C-ABI returns PYCD_ERR_ARG and the Python layer raises."""
import ctypes as C

import numpy as np
import pytest

from pycd_amd import batch as B, constants, rng, workloads
from pycd_amd._native import FamilyC, NativeError, lib

pytestmark = pytest.mark.gpu

PYCD_ERR_ARG = -1


@pytest.fixture(scope="module")
def small():
    w = workloads.corundum(3, t_final=2e-9)
    table = w.build_table()
    return B.Family(w.make_run(table), table)


def _err(bt):
    return lib().pycd_last_error(bt.ctx).decode()


def _family_struct(cells, n_unit, z=4):
    f = FamilyC()
    f.cells[0], f.cells[1], f.cells[2] = cells
    f.n_unit = n_unit
    f.n_sites = int(np.prod(cells)) * n_unit
    f.z = z
    f.max_z = z
    f.table_rows = n_unit
    return f  # table pointers left null: the sizes are checked first


@pytest.mark.parametrize("cells, n_unit, z, why", [
    ((1, 74, 74), 12, 4, "65536"),        # packed-site arithmetic
    ((256, 1, 1), 1, 4, "cells[0]"),      # cells per axis
    ((2, 2, 2), 4, 8, "z must be"),       # hop slots
])
def test_batch_family_limits(small, cells, n_unit, z, why):
    bt = B.KmcBatch([small], "f64", "f32")
    idx = C.c_int32(-1)
    f = _family_struct(cells, n_unit, z)
    rc = lib().pycd_batch_add_family(bt.h, C.byref(f), C.byref(idx))
    assert rc == PYCD_ERR_ARG
    assert why in _err(bt)
    bt.close()


def test_batch_setup_rejects_bad_trajectories(small):
    bt = B.KmcBatch([small], "f64", "f32")
    kT = 300.0 * constants.K2AUTEMP
    run = small.run                  # time grid in atomic units, 2 ns long
    good = dict(family=0, n_carriers=2, carrier_q=-1.0, field=np.zeros(3),
                kT=kT, time_interval=run.time_interval,
                n_path=int(run.t_final / run.time_interval) + 1, max_hops=10)
    occs, words = rng.seeded_starts(1, 1, [small.hop_sites], [2])
    for bad in (dict(n_carriers=65), dict(family=3), dict(time_interval=0.0),
                dict(n_path=0)):
        t = {**good, **bad}
        o = occs if t["n_carriers"] == 2 else [list(range(t["n_carriers"]))]
        with pytest.raises(NativeError):
            bt.setup([t], o, words)
    # the batch is still usable after the refusals
    bt.setup([good], occs, words)
    bt.run()
    assert bt.results()["hops"][0] == 10
    bt.close()

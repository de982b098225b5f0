"""pycd_kmc_advance_static (no allocation) against pycd_kmc_advance."""
import ctypes as C

import numpy as np
import pytest

from pycd_amd import _native, rng, workloads
from tests.kmc_helpers import native_problem, seeded_start

pytestmark = pytest.mark.gpu


def test_static_advance_equals_advance():
    """Two calls in a row, the second with fewer uniform pairs (5000, then
    1200): the allocation-free entry point gives the hop times, chosen
    processes and trajectory state of pycd_kmc_advance bit for bit; sizes
    beyond its fixed scratch and host buffers are refused before any
    kernel runs."""
    w = workloads.corundum(4, n_carriers=8, field_vpm=1e6, t_final=2e-8)
    table = w.build_table()
    run = w.make_run(table)
    gen, occ = seeded_start(run, 2, 8)
    n1, n2 = 5000, 1200
    uni = rng.UniformStream(gen.getstate()).take(2 * (n1 + n2))
    dev_u = _native.DeviceArray.from_host(uni)
    times = _native.DeviceArray(8 * n1, np.float64, (n1,))
    procs = _native.DeviceArray(4 * n1, np.int32, (n1,))
    ref = _native.KmcSession(native_problem(run, table))
    sta = _native.KmcSession(native_problem(run, table))
    try:
        ref.begin(occ)
        sta.begin(occ)
        off = 0
        for n in (n1, n2):
            rc_r, t_r, p_r = ref.advance(uni[2 * off:2 * (off + n)],
                                         record=True)
            rc_s, made = sta.advance_static(dev_u, n, dev_times=times,
                                            dev_procs=procs,
                                            offset_pairs=off)
            assert rc_s == rc_r
            assert made == len(t_r) == n
            assert np.array_equal(times.to_host()[:made], t_r)
            assert np.array_equal(procs.to_host()[:made], p_r)
            a, b = ref.read(), sta.read()
            for key in ("unwrapped", "energy", "delg_0", "drift"):
                assert np.array_equal(a[key], b[key]), key
            for key in ("occupancy", "t", "energy_now", "start", "hops"):
                assert a[key] == b[key], key
            off += n
        with pytest.raises(_native.NativeError, match="at most"):
            sta.advance_static(dev_u, (1 << 20) + 1)
        made = C.c_int64(0)
        rc = _native.lib().pycd_kmc_advance_static(
            sta.h, uni.ctypes.data, 10, 0, None, None, C.byref(made))
        assert rc == -1 and made.value == 0
        assert b"device buffers only" in _native.lib().pycd_last_error(
            sta.ctx)
    finally:
        ref.close()
        sta.close()
        for d in (dev_u, times, procs):
            d.free()

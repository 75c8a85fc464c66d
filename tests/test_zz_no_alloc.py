"""The allocation-free readout (dsim_signal_moment_nalloc) through the C ABI
against dsim_signal_moment on the same spins: two calls in a row, the second
on a smaller input, then the size limit."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESH = os.path.join(ROOT, "tests", "golden", "reference_inputs", "sim.fibers_1_r004_f09")
PC = np.float32(2 * np.pi * 42576.0 * 0.005)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fn():
    L = D.lib()
    f = L.dsim_signal_moment_nalloc
    f.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p,
                  C.c_uint32, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _case(n, lobes, ndir, seed):
    """A labelled moment-mode simulation; returns (sim, wsum, grads)."""
    m = O.parse_mesh(MESH, 6.0)
    sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"])
    sim = D.Simulation(sub, n, dt=0.005, std_dev=(0.03, 0.03, 0.01, 0.03),
                       t2=(80.0, 70.0, 7.5, 70.0), permeability=0.1, mode=D.MODE_PHYSICAL,
                       seed=seed)
    sim.init_uniform(0.9)
    sim.label(m["xmax"], m["ymax"], m["zmax"])
    sim.moment_reset()
    wsum = 0.0
    for w, k in lobes:
        sim.integrate_moment(w, k)
        wsum += w * k
    rng = np.random.default_rng(seed)
    g = (rng.standard_normal((ndir, 3)) * 2e-4).astype(np.float32)
    return sim, sub, np.float32(wsum), g


def _device_fields(sim):
    s = sim.get_state()
    mom, x0 = sim.get_moment()
    lab = s["comp"].astype(np.uint32) | (s["fiber"].astype(np.uint32) << 8)
    # device copies in the handle's field-major layout
    return (D.DeviceArray(s["mag"]), D.DeviceArray(mom.T.reshape(-1)),
            D.DeviceArray(x0.T.reshape(-1)), D.DeviceArray(lab))


def test_nalloc_readout_matches_signal_moment():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    f = _fn()
    # first call: 6000 spins, 96 directions, lobes that leave wsum = 100
    # (the x0 term is used); second call: 1500 spins, 7 directions
    for n, lobes, ndir, seed in ((6000, ((1.0, 300), (0.0, 50), (-1.0, 200)), 96, 5),
                                 (1500, ((1.0, 120), (-1.0, 120)), 7, 6)):
        _say(f"nalloc readout: {n} spins, {ndir} directions")
        sim, sub, wsum, g = _case(n, lobes, ndir, seed)
        ref = sim.signal_moment(g, PC)
        mag, mom, x0, lab = _device_fields(sim)
        gd = D.DeviceArray(g.reshape(-1))
        out = D.DeviceArray(np.full(ndir * 5, -1.0))
        rc = f(n, mag.ptr, mom.ptr, x0.ptr, float(wsum), lab.ptr, ndir, gd.ptr, float(PC),
               out.ptr, None)
        assert rc == 0, D.lib().dsim_last_error()
        D.check(D.lib().dsim_thread_sync())
        got = out.get(np.float64, (ndir, 5))
        # same kernels, same partial-sum order: the sums are identical; the
        # last division is done on the device here and on the host there
        np.testing.assert_allclose(got, ref, rtol=1e-15, atol=0)
        assert np.all(got[:, 0] > 0)
        for a in (mag, mom, x0, lab, gd, out):
            a.free()
        sim.close()
        sub.close()
    # beyond the fixed scratch: a clear error, nothing launched
    big = f(10, None, None, None, 0.0, None, 257, None, 1.0, None, None)
    assert big == -1
    assert b"at most 256" in D.lib().dsim_last_error()

"""ctypes binding of the waveform phases (include/dsim_phase.h).

Like dsim_amd.lib it raises when libdsim.so is missing and never falls back to
a CPU implementation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as D

_ready = False


def _lib():
    global _ready
    L = D.lib()
    if not _ready:
        vp, fp, dp = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double)
        sigs = {
            "dsim_waveforms_create": [vp, C.c_uint32, C.c_uint32, fp, C.c_float, C.c_uint32,
                                      C.POINTER(vp)],
            "dsim_waveforms_destroy": [vp],
            "dsim_waveforms_set_buffer": [vp, C.c_uint32],
            "dsim_waveforms_run": [vp, C.c_uint32, fp],
            "dsim_waveforms_signal": [vp, dp],
            "dsim_waveforms_get_phase": [vp, fp, fp],
            "dsim_waveforms_timing": [vp, dp, dp],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _ready = True
    return L


def _check(rc):
    if rc != D.OK:
        raise D.DsimError(rc, D.lib().dsim_last_error().decode())


class Waveforms:
    """Phases of n_wave arbitrary gradient waveforms over one walk of `sim`.
    grads: [n_steps, n_wave, 3] gradScaled values."""

    def __init__(self, sim, grads, phase_constant, chunk_steps=256, buffer_steps=None):
        g = np.ascontiguousarray(grads, np.float32)
        self.n_steps, self.n_wave = g.shape[0], g.shape[1]
        self.n = sim.n
        self.sim = sim
        self.h = C.c_void_p()
        _check(_lib().dsim_waveforms_create(sim.h, self.n_wave, self.n_steps,
                                            g.ctypes.data_as(C.POINTER(C.c_float)),
                                            float(phase_constant), chunk_steps, C.byref(self.h)))
        if buffer_steps is not None:
            _check(_lib().dsim_waveforms_set_buffer(self.h, int(buffer_steps)))

    def close(self):
        if self.h:
            _check(_lib().dsim_waveforms_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, steps, trajectory=False):
        tr = np.zeros((steps, 3, self.n), np.float32) if trajectory else None
        _check(_lib().dsim_waveforms_run(
            self.h, steps, None if tr is None else tr.ctypes.data_as(C.POINTER(C.c_float))))
        return tr

    def signal(self):
        out = np.zeros((self.n_wave, 5), np.float64)
        _check(_lib().dsim_waveforms_signal(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def phase(self):
        """(phase, compensation), each [n_spins, n_wave], without the x0 term."""
        p = np.zeros((self.n, self.n_wave), np.float32)
        c = np.zeros((self.n, self.n_wave), np.float32)
        _check(_lib().dsim_waveforms_get_phase(self.h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                               c.ctypes.data_as(C.POINTER(C.c_float))))
        return p, c

    def timing(self):
        a, b = C.c_double(), C.c_double()
        _check(_lib().dsim_waveforms_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value


def random_waveforms(n_wave, n_steps, seed, g_max=80.0, hold=2):
    """Piecewise-constant gradients [n_steps, n_wave, 3] in mT/m: every `hold`
    steps each waveform takes a new vector, direction uniform on the sphere and
    magnitude uniform in [0, g_max]."""
    rng = np.random.default_rng(seed)
    segs = (n_steps + hold - 1) // hold
    v = rng.normal(size=(segs, n_wave, 3))
    v /= np.linalg.norm(v, axis=2, keepdims=True)
    v *= rng.uniform(0.0, g_max, (segs, n_wave, 1))
    return np.repeat(v, hold, axis=0)[:n_steps].astype(np.float32)

"""ctypes binding of the dsim server (include/dsim_serve.h).

A Server keeps Substrates resident on the device and runs simulation
requests in batched rounds.  Like dsim_amd.lib it raises when libdsim.so is
missing and never falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as D


class Request(C.Structure):
    _fields_ = [("substrate", C.c_uint32), ("n_spins", C.c_uint32), ("seed", C.c_uint64),
                ("dt", C.c_float), ("std_dev", C.c_float * 4), ("t2", C.c_float * 4),
                ("permeability", C.c_float), ("boundary", C.c_uint32), ("start_box", C.c_float),
                ("label_extent", C.c_float * 3), ("n_lobes", C.c_uint32),
                ("lobe_weight", C.POINTER(C.c_float)), ("lobe_steps", C.POINTER(C.c_uint32)),
                ("n_lines", C.c_uint32), ("grads", C.POINTER(C.c_float)),
                ("phase_constant", C.c_float), ("voxel_lo", C.c_float * 3),
                ("voxel_hi", C.c_float * 3)]


class RequestStats(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("spin_steps", C.c_uint64), ("rounds", C.c_uint32), ("ambiguous", C.c_uint32)]


_ready = False


def _lib():
    global _ready
    L = D.lib()
    if not _ready:
        vp = C.c_void_p
        sigs = {
            "dsim_server_create": [C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(vp)],
            "dsim_server_destroy": [vp],
            "dsim_server_add_substrate": [vp, vp, C.POINTER(C.c_uint32)],
            "dsim_server_submit": [vp, C.POINTER(Request), C.POINTER(C.c_uint64)],
            "dsim_server_round": [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
            "dsim_server_set_policy": [vp, C.c_uint32],
            "dsim_server_start": [vp],
            "dsim_server_stop": [vp],
            "dsim_server_wait": [vp, C.c_uint64, C.c_double],
            "dsim_server_result": [vp, C.c_uint64, C.POINTER(C.c_double), C.POINTER(RequestStats)],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.dsim_server_clock.argtypes = [vp]
        L.dsim_server_clock.restype = C.c_double
        _ready = True
    return L


def _check(rc):
    if rc != D.OK:
        raise D.DsimError(rc, D.lib().dsim_last_error().decode())


def pgse_lobes(L, B, R=0):
    """Gradient lobes of one PGSE line (diffusion.cpp:564-585): +G for L steps,
    0 until B, -G for L steps, 0 for R steps."""
    return [(1.0, L), (0.0, B - L), (-1.0, L), (0.0, R)]


class Server:
    def __init__(self, hbm_budget, round_steps=256, max_active_spins=0, small_first_blocks=0):
        """small_first_blocks: see dsim_server_set_policy (0: every prepared
        request in every round)."""
        self.h = C.c_void_p()
        _check(_lib().dsim_server_create(int(hbm_budget), round_steps, max_active_spins,
                                         C.byref(self.h)))
        if small_first_blocks:
            _check(_lib().dsim_server_set_policy(self.h, int(small_first_blocks)))
        self._subs = []
        self._lines = {}

    def close(self):
        if self.h:
            _check(_lib().dsim_server_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_substrate(self, sub):
        i = C.c_uint32()
        _check(_lib().dsim_server_add_substrate(self.h, sub.h, C.byref(i)))
        self._subs.append(sub)  # keep the substrate alive as long as the server
        return i.value

    def submit(self, substrate, n_spins, *, seed, dt, std_dev, t2, permeability, lobes, grads,
               phase_constant, boundary=D.BOUNDARY_PERIODIC, start_box=1.0,
               label_extent=(1.1, 1.1, 1.1), voxel=None):
        q = Request()
        q.substrate, q.n_spins, q.seed = substrate, n_spins, seed
        q.dt = dt
        q.std_dev[:] = [float(x) for x in np.broadcast_to(std_dev, 4)]
        q.t2[:] = [float(x) for x in np.broadcast_to(t2, 4)]
        q.permeability, q.boundary, q.start_box = permeability, boundary, start_box
        q.label_extent[:] = list(label_extent)
        w = np.ascontiguousarray([float(a) for a, _ in lobes], np.float32)
        n = np.ascontiguousarray([int(b) for _, b in lobes], np.uint32)
        g = np.ascontiguousarray(np.asarray(grads, np.float32).reshape(-1, 3))
        q.n_lobes = len(w)
        q.lobe_weight = w.ctypes.data_as(C.POINTER(C.c_float))
        q.lobe_steps = n.ctypes.data_as(C.POINTER(C.c_uint32))
        q.n_lines = len(g)
        q.grads = g.ctypes.data_as(C.POINTER(C.c_float))
        q.phase_constant = phase_constant
        if voxel is not None:  # ((lo x, y, z), (hi x, y, z)): non-cubic periodic voxel
            q.voxel_lo[:] = [float(v) for v in voxel[0]]
            q.voxel_hi[:] = [float(v) for v in voxel[1]]
        t = C.c_uint64()
        _check(_lib().dsim_server_submit(self.h, C.byref(q), C.byref(t)))
        self._lines[t.value] = len(g)
        return t.value

    def round(self):
        a, q = C.c_uint32(), C.c_uint32()
        _check(_lib().dsim_server_round(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def run_until_idle(self):
        while True:
            a, q = self.round()
            if a == 0 and q == 0:
                return

    def start(self):
        _check(_lib().dsim_server_start(self.h))

    def stop(self):
        _check(_lib().dsim_server_stop(self.h))

    def wait(self, ticket, timeout_s=-1.0):
        _check(_lib().dsim_server_wait(self.h, ticket, timeout_s))

    def result(self, ticket):
        """(signals [n_lines, 5], stats dict)."""
        out = np.zeros((self._lines[ticket], 5), np.float64)
        st = RequestStats()
        _check(_lib().dsim_server_result(self.h, ticket, out.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.byref(st)))
        return out, {k: getattr(st, k) for k, _ in RequestStats._fields_}

    def clock(self):
        return _lib().dsim_server_clock(self.h)

"""ctypes binding of libdsim.so (the C ABI declared in include/dsim.h).

The product path is the HIP library; if it is missing this module raises,
it never substitutes a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsim.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "dsim.h")

OK = 0
ERR = {-1: "ARG", -2: "HIP", -3: "NOMEM", -4: "STATE", -5: "OVERFLOW", -6: "IO", -7: "FORMAT",
       -8: "NODEV"}
NO_TRIANGLE = 0xFFFFFFFF
NO_FIBER = 65535
GRID_REFERENCE, GRID_AABB, GRID_TRIBOX = 0, 1, 2
# DSIM_STAT_* of include/dsim.h, in index order
WALK_STATS = ("steps", "trips", "search_lanes", "cells", "plane_tests", "plane_passes",
              "tri_tests", "tri_passes", "fp64_tests", "hits", "shortcuts", "deferred",
              "waves", "live_lanes")
MODE_REFERENCE, MODE_PHYSICAL = 0, 1
BOUNDARY_CLAMP, BOUNDARY_PERIODIC = 0, 1
RNG_MWC, RNG_PHILOX = 1, 2


class DsimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dsim error {code} ({ERR.get(code, '?')}): {msg}")
        self.code = code


class SpinData(C.Structure):
    """spinData, dSim3/src/dSimDataTypes.h:9-24."""
    _fields_ = [("signalMagnitude", C.c_float), ("signalPhase", C.c_float),
                ("compartmentType", C.c_uint8), ("insideFiber", C.c_uint16)]


SPIN_DTYPE = np.dtype({"names": ["signalMagnitude", "signalPhase", "compartmentType", "insideFiber"],
                       "formats": [np.float32, np.float32, np.uint8, np.uint16],
                       "offsets": [0, 4, 8, 10], "itemsize": 12})


class Mesh(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("vertices", C.POINTER(C.c_float)),
                ("n_triangles", C.c_uint32), ("triangles", C.POINTER(C.c_uint32)),
                ("tri_info", C.POINTER(C.c_uint32))]


class GridParams(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("dims", C.c_uint32 * 3), ("lo", C.c_float * 3),
                ("cell", C.c_float * 3), ("eps", C.c_float), ("clearance_rings", C.c_uint32)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_float), ("std_dev", C.c_float * 4), ("t2", C.c_float * 4),
                ("permeability", C.c_float), ("reflection_type", C.c_uint32), ("mode", C.c_uint32),
                ("boundary", C.c_uint32), ("rng", C.c_uint32), ("seed", C.c_uint64),
                ("spin_radius", C.c_float), ("sort_interval", C.c_uint32),
                ("steps_per_launch", C.c_uint32), ("voxel_lo", C.c_float * 3),
                ("voxel_hi", C.c_float * 3), ("fixed_blocks", C.c_uint32)]


_lib = None
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def exported_names():
    """Function names declared in include/*.h."""
    inc = os.path.dirname(HEADER)
    names = set()
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            src = open(os.path.join(inc, h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dsim_\w+)\s*\(", src, re.M))
    return sorted(names)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dsim_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.dsim_last_error.restype = C.c_char_p
        L.dsim_abi_version.restype = C.c_int
        sigs = {
            "dsim_device_count": [C.POINTER(C.c_int)],
            "dsim_alloc_array": [C.POINTER(_vp), C.c_size_t],
            "dsim_free_array": [_vp],
            "dsim_thread_sync": [],
            "dsim_copy_from_device": [_vp, _vp, C.c_size_t],
            "dsim_copy_to_device": [_vp, _vp, C.c_size_t, C.c_size_t],
            "dsim_substrate_create": [C.POINTER(Mesh), C.POINTER(GridParams), C.POINTER(_vp)],
            "dsim_substrate_create_device": [C.POINTER(Mesh), C.POINTER(GridParams),
                                             C.POINTER(_vp)],
            "dsim_substrate_destroy": [_vp],
            "dsim_substrate_grid_info": [_vp, _up, _up, _up],
            "dsim_substrate_get_grid": [_vp, _up, _up],
            "dsim_substrate_get_helpers": [_vp, _fp],
            "dsim_substrate_bytes": [_vp, C.POINTER(C.c_uint64)],
            "dsim_substrate_label_axis": [_vp, _up],
            "dsim_sim_create": [_vp, C.POINTER(Params), C.c_uint32, C.POINTER(_vp)],
            "dsim_sim_destroy": [_vp],
            "dsim_set_state_aos": [_vp, _fp, _vp, _up],
            "dsim_get_state_aos": [_vp, _fp, _vp, _up],
            "dsim_set_step": [_vp, C.c_uint64],
            "dsim_get_step": [_vp, C.POINTER(C.c_uint64)],
            "dsim_init_spins_uniform": [_vp, C.c_float],
            "dsim_init_spins_box": [_vp, _fp, _fp],
            "dsim_label_spins": [_vp, C.c_float, C.c_float, C.c_float, C.c_uint32, _up],
            "dsim_integrate": [_vp, _fp, C.c_float, C.c_uint32],
            "dsim_integrate_injected": [_vp, _fp, C.c_float, _fp, C.c_uint32, _up],
            "dsim_signal": [_vp, _dp],
            "dsim_moment_reset": [_vp],
            "dsim_integrate_moment": [_vp, C.c_float, C.c_uint32],
            "dsim_integrate_moment_injected": [_vp, C.c_float, _fp, C.c_uint32, _up],
            "dsim_signal_moment": [_vp, C.c_uint32, _fp, C.c_float, _dp],
            "dsim_get_moment": [_vp, _fp, _fp],
            "dsim_sync": [_vp],
            "dsim_step_timing": [_vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
            "dsim_step_timing_reset": [_vp],
            "dsim_walk_stats_enable": [_vp, C.c_int],
            "dsim_walk_stats_read": [_vp, C.POINTER(C.c_uint64)],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise DsimError(rc, lib().dsim_last_error().decode(errors="replace"))
    return rc


def device_count():
    n = C.c_int(0)
    check(lib().dsim_device_count(C.byref(n)))
    return n.value


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class DeviceArray:
    """A device copy of a numpy array through the reference-style support
    entries (dsim_alloc_array, dsim_copy_to_device, dsim_copy_from_device)."""

    def __init__(self, a=None, *, nbytes=None):
        self.nbytes = int(a.nbytes if a is not None else nbytes)
        self.ptr = _vp()
        check(lib().dsim_alloc_array(C.byref(self.ptr), self.nbytes))
        if a is not None:
            a = np.ascontiguousarray(a)
            check(lib().dsim_copy_to_device(self.ptr, a.ctypes.data_as(_vp), 0, self.nbytes))

    def get(self, dtype, shape=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        check(lib().dsim_copy_from_device(out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out if shape is None else out.reshape(shape)

    def free(self):
        if self.ptr:
            check(lib().dsim_free_array(self.ptr))
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Substrate:
    """Triangle mesh + uniform grid resident on the device.  device=True builds
    the records, grid and clearance on the GPU (dsim_substrate_create_device);
    the arrays equal the host build's."""

    def __init__(self, vertices, triangles, tri_info, *, grid="reference", n_cubes=50, dims=None,
                 lo=None, cell=None, eps=0.0, device=False, clearance_rings=0):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        self.tri_info = np.ascontiguousarray(tri_info, np.uint32).reshape(-1, 3)
        m = Mesh(len(self.vertices), _p(self.vertices, C.c_float), len(self.triangles),
                 _p(self.triangles, C.c_uint32), _p(self.tri_info, C.c_uint32))
        g = GridParams()
        if grid == "reference":
            g.kind = GRID_REFERENCE
            g.dims = (C.c_uint32 * 3)(n_cubes, n_cubes, n_cubes)
        else:  # "aabb" or "tribox" (exact triangle/box membership)
            g.kind = GRID_TRIBOX if grid == "tribox" else GRID_AABB
            g.dims = (C.c_uint32 * 3)(*[int(d) for d in dims])
            g.lo = (C.c_float * 3)(*[float(x) for x in lo])
            g.cell = (C.c_float * 3)(*[float(x) for x in cell])
            g.eps = float(eps)
            g.clearance_rings = int(clearance_rings)
        h = _vp()
        create = lib().dsim_substrate_create_device if device else lib().dsim_substrate_create
        check(create(C.byref(m), C.byref(g), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(lib().dsim_substrate_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            if self.h:
                lib().dsim_substrate_destroy(self.h)
        except Exception:
            pass

    def grid(self):
        nc, ne, mx = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().dsim_substrate_grid_info(self.h, C.byref(nc), C.byref(ne), C.byref(mx)))
        start = np.zeros(nc.value + 1, np.uint32)
        tri = np.zeros(max(1, ne.value), np.uint32)
        check(lib().dsim_substrate_get_grid(self.h, _p(start, C.c_uint32), _p(tri, C.c_uint32)))
        return start, tri[: ne.value], mx.value

    def helpers(self):
        h = np.zeros((len(self.triangles), 12), np.float32)
        check(lib().dsim_substrate_get_helpers(self.h, _p(h, C.c_float)))
        return h

    def clearance(self):
        nc, ne, mx = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().dsim_substrate_grid_info(self.h, C.byref(nc), C.byref(ne), C.byref(mx)))
        out = np.zeros(nc.value, np.float32)
        L = lib()
        L.dsim_substrate_get_clearance.argtypes = [_vp, _fp]
        L.dsim_substrate_get_clearance.restype = C.c_int
        check(L.dsim_substrate_get_clearance(self.h, _p(out, C.c_float)))
        return out

    def label_axis(self):
        """Axis of the physical mode's labelling rays: 0 x, 1 y, 2 z."""
        a = C.c_uint32()
        check(lib().dsim_substrate_label_axis(self.h, C.byref(a)))
        return a.value

    def nbytes(self):
        b = C.c_uint64()
        check(lib().dsim_substrate_bytes(self.h, C.byref(b)))
        return b.value


class Simulation:
    """One set of spins on one substrate (the device half of SpinSystem)."""

    def __init__(self, substrate: Substrate, n, *, dt, std_dev, t2=(np.inf,) * 4, permeability=0.0,
                 reflection_type=1, mode=MODE_REFERENCE, boundary=BOUNDARY_CLAMP, rng=RNG_PHILOX,
                 seed=1, spin_radius=0.005, sort_interval=0, steps_per_launch=0, voxel=None,
                 fixed_blocks=False):
        self.substrate = substrate
        self.n = int(n)
        p = Params()
        p.dt = float(dt)
        p.std_dev = (C.c_float * 4)(*[float(x) for x in std_dev])
        p.t2 = (C.c_float * 4)(*[float(x) for x in t2])
        p.permeability = float(permeability)
        p.reflection_type = reflection_type
        p.mode = mode
        p.boundary = boundary
        p.rng = rng
        p.seed = int(seed) & ((1 << 64) - 1)
        p.spin_radius = float(spin_radius)
        p.sort_interval = int(sort_interval)
        p.steps_per_launch = int(steps_per_launch)
        p.fixed_blocks = int(bool(fixed_blocks))
        if voxel is not None:  # ((lo x, y, z), (hi x, y, z)) of a non-cubic periodic voxel
            p.voxel_lo = (C.c_float * 3)(*[float(v) for v in voxel[0]])
            p.voxel_hi = (C.c_float * 3)(*[float(v) for v in voxel[1]])
        self.params = p
        h = _vp()
        check(lib().dsim_sim_create(substrate.h, C.byref(p), self.n, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(lib().dsim_sim_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            if self.h:
                lib().dsim_sim_destroy(self.h)
        except Exception:
            pass

    # -- state ------------------------------------------------------------
    def set_state(self, pos=None, comp=None, fiber=None, mag=None, phase=None, seeds=None):
        P = None if pos is None else np.ascontiguousarray(pos, np.float32).reshape(-1)
        S = None
        if comp is not None or mag is not None or phase is not None or fiber is not None:
            S = np.zeros(self.n, SPIN_DTYPE)
            S["signalMagnitude"] = 1.0 if mag is None else mag
            S["signalPhase"] = 0.0 if phase is None else phase
            S["compartmentType"] = 0 if comp is None else comp
            S["insideFiber"] = NO_FIBER if fiber is None else fiber
        sd = None if seeds is None else np.ascontiguousarray(seeds, np.uint32).reshape(-1)
        check(lib().dsim_set_state_aos(self.h, _p(P, C.c_float),
                                       None if S is None else S.ctypes.data_as(_vp),
                                       _p(sd, C.c_uint32)))

    def get_state(self, seeds=False):
        P = np.zeros(self.n * 3, np.float32)
        S = np.zeros(self.n, SPIN_DTYPE)
        sd = np.zeros(self.n * 2, np.uint32) if seeds else None
        check(lib().dsim_get_state_aos(self.h, _p(P, C.c_float), S.ctypes.data_as(_vp),
                                       _p(sd, C.c_uint32)))
        out = dict(pos=P.reshape(-1, 3), mag=S["signalMagnitude"].copy(),
                   phase=S["signalPhase"].copy(), comp=S["compartmentType"].copy(),
                   fiber=S["insideFiber"].copy())
        if seeds:
            out["seeds"] = sd.reshape(-1, 2)
        return out

    def set_step(self, step):
        check(lib().dsim_set_step(self.h, int(step)))

    def init_uniform(self, start_box):
        check(lib().dsim_init_spins_uniform(self.h, float(start_box)))

    def init_box(self, lo, hi):
        a = np.ascontiguousarray(lo, np.float32)
        b = np.ascontiguousarray(hi, np.float32)
        check(lib().dsim_init_spins_box(self.h, _p(a, C.c_float), _p(b, C.c_float)))

    def label(self, xmax=1.0, ymax=1.0, zmax=1.0, max_tries=20):
        amb = C.c_uint32()
        check(lib().dsim_label_spins(self.h, xmax, ymax, zmax, max_tries, C.byref(amb)))
        return amb.value

    # -- stepping ---------------------------------------------------------
    def integrate(self, grad, phase_constant, iters):
        g = np.ascontiguousarray(grad, np.float32)
        check(lib().dsim_integrate(self.h, _p(g, C.c_float), float(phase_constant), int(iters)))

    def integrate_injected(self, grad, phase_constant, n3u, trace=False):
        g = np.ascontiguousarray(grad, np.float32)
        inj = np.ascontiguousarray(n3u, np.float32)
        iters = inj.shape[0]
        assert inj.shape == (iters, self.n, 4)
        tr = np.zeros((iters, self.n, 5), np.uint32) if trace else None
        check(lib().dsim_integrate_injected(self.h, _p(g, C.c_float), float(phase_constant),
                                            _p(inj, C.c_float), iters, _p(tr, C.c_uint32)))
        return tr

    def signal(self):
        out = np.zeros(5, np.float64)
        check(lib().dsim_signal(self.h, _p(out, C.c_double)))
        return out

    def moment_reset(self):
        check(lib().dsim_moment_reset(self.h))

    def integrate_moment(self, weight, iters):
        check(lib().dsim_integrate_moment(self.h, float(weight), int(iters)))

    def integrate_moment_injected(self, weight, n3u, trace=False):
        inj = np.ascontiguousarray(n3u, np.float32)
        iters = inj.shape[0]
        assert inj.shape == (iters, self.n, 4)
        tr = np.zeros((iters, self.n, 5), np.uint32) if trace else None
        check(lib().dsim_integrate_moment_injected(self.h, float(weight), _p(inj, C.c_float),
                                                   iters, _p(tr, C.c_uint32)))
        return tr

    def signal_moment(self, grads_scaled, phase_constant):
        g = np.ascontiguousarray(grads_scaled, np.float32).reshape(-1, 3)
        out = np.zeros((len(g), 5), np.float64)
        check(lib().dsim_signal_moment(self.h, len(g), _p(g, C.c_float), float(phase_constant),
                                       _p(out, C.c_double)))
        return out

    def get_moment(self):
        m = np.zeros((self.n, 3), np.float32)
        x0 = np.zeros((self.n, 3), np.float32)
        check(lib().dsim_get_moment(self.h, _p(m, C.c_float), _p(x0, C.c_float)))
        return m, x0

    def sync(self):
        check(lib().dsim_sync(self.h))

    def step_timing(self):
        ms = C.c_double()
        n = C.c_uint64()
        check(lib().dsim_step_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def step_timing_reset(self):
        check(lib().dsim_step_timing_reset(self.h))

    def walk_stats(self, enable=None):
        """Work counters of the physical-mode walk (a dict keyed by the
        WALK_STATS names).  walk_stats(True) zeroes them and switches the
        counting build on; walk_stats(False) switches it off; walk_stats()
        reads them."""
        if enable is not None:
            check(lib().dsim_walk_stats_enable(self.h, int(bool(enable))))
            return None
        a = (C.c_uint64 * 16)()
        check(lib().dsim_walk_stats_read(self.h, a))
        return {k: int(a[i]) for i, k in enumerate(WALK_STATS)}


# --------------------------------------------------------------------------- #
# I/O (include/dsim_io.h)
# --------------------------------------------------------------------------- #
PATH_MAX = 1024


class Config(C.Structure):
    _fields_ = [("stepsPerUpdate", C.c_int32), ("useGpu", C.c_int32), ("useDisplay", C.c_int32),
                ("writeOut", C.c_int32), ("numSpins", C.c_uint32),
                ("gyroMagneticRatio", C.c_float), ("timeStep", C.c_float),
                ("extraAdc", C.c_float), ("intraAdc", C.c_float), ("myelinAdc", C.c_float),
                ("extraT2", C.c_float), ("intraT2", C.c_float), ("myelinT2", C.c_float),
                ("spaceScale", C.c_float), ("startBox", C.c_float), ("permeability", C.c_float),
                ("searchMethod", C.c_uint32), ("outFile", C.c_char * PATH_MAX),
                ("gradsFile", C.c_char * PATH_MAX), ("fiberFile", C.c_char * PATH_MAX),
                ("configFile", C.c_char * PATH_MAX), ("seed", C.c_uint64)]


class MeshInfo(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("n_triangles", C.c_uint32), ("n_axon", C.c_uint32),
                ("n_myelin", C.c_uint32), ("n_glia", C.c_uint32), ("n_fibers", C.c_uint32),
                ("max_triangles_on_surface", C.c_uint32), ("xmax", C.c_float),
                ("ymax", C.c_float), ("zmax", C.c_float)]


class TensorFit(C.Structure):
    _fields_ = [("tensor", (C.c_double * 3) * 3), ("eigvec", (C.c_double * 3) * 3),
                ("eigval", C.c_float * 3), ("md", C.c_float), ("fa", C.c_float),
                ("rd", C.c_float)]


_io_ready = False


def _io():
    global _io_ready
    L = lib()
    if not _io_ready:
        L.dsim_io_last_error.restype = C.c_char_p
        sigs = {
            "dsim_config_defaults": [C.POINTER(Config)],
            "dsim_config_read": [C.c_char_p, C.POINTER(Config)],
            "dsim_config_apply_args": [C.POINTER(Config), C.c_int, C.POINTER(C.c_char_p)],
            "dsim_grads_read": [C.c_char_p, _fp, C.c_uint32, _up],
            "dsim_mesh_read": [C.c_char_p, C.c_float, C.POINTER(_vp)],
            "dsim_mesh_view": [_vp, C.POINTER(Mesh), C.POINTER(MeshInfo)],
            "dsim_mesh_free": [_vp],
            "dsim_mesh_write": [C.c_char_p, C.POINTER(Mesh), C.c_float],
            "dsim_pgse_schedule": [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                                   C.POINTER(C.c_int32), C.c_uint32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)],
            "dsim_tensor_fit_compute": [C.c_uint32, _fp, _fp, C.POINTER(TensorFit)],
            "dsim_write_m": [C.c_char_p, C.POINTER(Config), C.c_uint32, _fp, _fp,
                             C.POINTER(TensorFit)],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _io_ready = True
    return L


def io_check(rc):
    if rc != OK:
        raise DsimError(rc, _io().dsim_io_last_error().decode(errors="replace"))
    return rc


def read_config(path=None, argv=()):
    c = Config()
    io_check(_io().dsim_config_defaults(C.byref(c)))
    if path is not None:
        io_check(_io().dsim_config_read(str(path).encode(), C.byref(c)))
    if argv:
        arr = (C.c_char_p * (len(argv) + 1))(b"dsim", *[a.encode() for a in argv])
        io_check(_io().dsim_config_apply_args(C.byref(c), len(argv) + 1, arr))
    return c


def read_grads(path, cap=1 << 16):
    rows = np.zeros(cap * 6, np.float32)
    n = C.c_uint32()
    io_check(_io().dsim_grads_read(str(path).encode(), _p(rows, C.c_float), cap, C.byref(n)))
    return rows[: n.value * 6].reshape(-1, 6)


def read_mesh(path, space_scale):
    h = _vp()
    io_check(_io().dsim_mesh_read(str(path).encode(), float(space_scale), C.byref(h)))
    try:
        m, info = Mesh(), MeshInfo()
        io_check(_io().dsim_mesh_view(h, C.byref(m), C.byref(info)))
        nv, nt = info.n_vertices, info.n_triangles
        out = {k: getattr(info, k) for k, _ in MeshInfo._fields_}
        out["vertices"] = (np.ctypeslib.as_array(m.vertices, (nv * 3,)).reshape(-1, 3).copy()
                           if nv else np.zeros((0, 3), np.float32))
        out["triangles"] = (np.ctypeslib.as_array(m.triangles, (nt * 3,)).reshape(-1, 3).copy()
                            if nt else np.zeros((0, 3), np.uint32))
        out["tri_info"] = (np.ctypeslib.as_array(m.tri_info, (nt * 3,)).reshape(-1, 3).copy()
                           if nt else np.zeros((0, 3), np.uint32))
        return out
    finally:
        _io().dsim_mesh_free(h)


def write_mesh(path, vertices, triangles, tri_info, space_scale):
    V = np.ascontiguousarray(vertices, np.float32)
    T = np.ascontiguousarray(triangles, np.uint32)
    I = np.ascontiguousarray(tri_info, np.uint32)
    m = Mesh(len(V), _p(V, C.c_float), len(T), _p(T, C.c_uint32), _p(I, C.c_uint32))
    io_check(_io().dsim_mesh_write(str(path).encode(), C.byref(m), float(space_scale)))


def pgse_schedule(ldelta, bdelta, readout, time_step, steps_per_update, cap=1 << 22):
    sign = np.zeros(cap, np.int32)
    L, B, R, ro = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    io_check(_io().dsim_pgse_schedule(ldelta, bdelta, readout, time_step, steps_per_update,
                                      sign.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(L),
                                      C.byref(B), C.byref(R), C.byref(ro)))
    n = L.value + B.value + R.value
    return sign[:n].copy(), (L.value, B.value, R.value, ro.value)


def tensor_fit(rows, signal):
    R = np.ascontiguousarray(rows, np.float32)
    S_ = np.ascontiguousarray(signal, np.float32)
    t = TensorFit()
    io_check(_io().dsim_tensor_fit_compute(len(R), _p(R, C.c_float), _p(S_, C.c_float),
                                           C.byref(t)))
    return dict(tensor=np.array([list(r) for r in t.tensor]), eigvec=np.array([list(r) for r in t.eigvec]),
                eigval=np.array(t.eigval[:]), md=t.md, fa=t.fa, rd=t.rd)


def write_m(path, config, rows, signals, fits):
    R = np.ascontiguousarray(rows, np.float32)
    Sg = np.ascontiguousarray(signals, np.float32)
    arr = (TensorFit * len(fits))()
    for k, f in enumerate(fits):
        for i in range(3):
            for j in range(3):
                arr[k].tensor[i][j] = f["tensor"][i][j]
                arr[k].eigvec[i][j] = f["eigvec"][i][j]
        for i in range(3):
            arr[k].eigval[i] = f["eigval"][i]
        arr[k].md, arr[k].fa, arr[k].rd = f["md"], f["fa"], f["rd"]
    io_check(_io().dsim_write_m(str(path).encode(), C.byref(config), len(R), _p(R, C.c_float),
                                _p(Sg, C.c_float), arr))

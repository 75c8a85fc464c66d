"""ctypes front end of the dsim oracle (plain-C restatement of dSim3's CPU path).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsim_oracle.so")
NONE = 0xFFFFFFFF
FIBER_NONE = 65535

RNG_INJECT, RNG_MWC, RNG_PHILOX = 0, 1, 2


def build() -> None:
    subprocess.check_call(["sh", os.path.join(HERE, "build.sh")])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class F3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Coll(C.Structure):
    _fields_ = [("type", C.c_uint32), ("pt", F3), ("idx", C.c_uint32), ("dist_sq", C.c_float)]


class Hits(C.Structure):
    _fields_ = [("first", C.c_uint32), ("last", C.c_uint32), ("ncoll", C.c_uint32)]


_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)


class WorldC(C.Structure):
    _fields_ = [
        ("n_tri", C.c_uint32), ("helpers", _fp), ("vertices", _fp), ("tri_info", _up),
        ("trgls", _up), ("tri_mt", _fp),
        ("dims", C.c_uint32 * 3), ("lo", C.c_float * 3), ("cell", C.c_float * 3),
        ("inv_cell", C.c_float * 3), ("ref_cell_math", C.c_uint32), ("max_per_cell", C.c_uint32),
        ("cell_count", _up), ("cell_list", _up),
        ("std_dev", C.c_float * 4), ("t2", C.c_float * 4), ("decay", C.c_float * 4),
        ("dt", C.c_float), ("permeability", C.c_float), ("reflection_type", C.c_uint32),
        ("spin_radius", C.c_float), ("boundary", C.c_uint32), ("test", C.c_uint32),
        ("traversal", C.c_uint32), ("transitions", C.c_uint32),
        ("fiber_has_myelin", _u8p), ("n_fibers", C.c_uint32),
        ("wrap_lo", C.c_float * 3), ("wrap_hi", C.c_float * 3),
    ]


class StateC(C.Structure):
    _fields_ = [("n", C.c_uint32), ("pos", _fp), ("mag", _fp), ("phase", _fp), ("comp", _u8p),
                ("fiber", _u16p), ("seed", _up), ("disp", _fp), ("moment", _fp), ("moment_c", _fp),
                ("wrap", C.POINTER(C.c_int32))]


class StepArgsC(C.Structure):
    _fields_ = [("rng", C.c_uint32), ("philox_key", C.c_uint32 * 2), ("step0", C.c_uint64),
                ("inject", _fp), ("grad", F3), ("phase_constant", C.c_float), ("weight", C.c_float),
                ("trace", _up), ("n_threads", C.c_int), ("stats", C.POINTER(C.c_uint64))]


class MeshC(C.Structure):
    _fields_ = [("n_vert", C.c_uint32), ("n_tri", C.c_uint32), ("n_axon", C.c_uint32),
                ("n_myelin", C.c_uint32), ("n_glia", C.c_uint32), ("n_fibers", C.c_uint32),
                ("max_tri_on_surface", C.c_uint32), ("xmax", C.c_float), ("ymax", C.c_float),
                ("zmax", C.c_float), ("vertices", _fp), ("trgls", _up), ("tri_info", _up)]


def _declare(L):
    P = C.POINTER
    L.orc_mwc_next.argtypes = [_up]; L.orc_mwc_next.restype = C.c_uint32
    L.orc_box_muller.argtypes = [_fp, _fp]
    L.orc_randn.argtypes = [_up, _fp, _fp, _fp, _fp]
    L.orc_philox4x32_10.argtypes = [_up, _up, _up]
    L.orc_u32_to_unit.argtypes = [C.c_uint32]; L.orc_u32_to_unit.restype = C.c_float
    L.orc_philox_step_normals.argtypes = [_up, C.c_uint32, C.c_uint64, _fp]
    L.orc_philox_step_uniform.argtypes = [_up, C.c_uint32, C.c_uint64]
    L.orc_philox_step_uniform.restype = C.c_float
    L.orc_tri_helpers.argtypes = [_fp, _up, C.c_uint32, _fp]
    L.orc_tri_mt_records.argtypes = [_fp, _up, C.c_uint32, _fp]
    L.orc_cube_pos.argtypes = [P(WorldC), F3, _up]
    L.orc_cube_hash.argtypes = [P(WorldC), _up]; L.orc_cube_hash.restype = C.c_uint32
    L.orc_point_in_cube.argtypes = [F3, F3, F3]; L.orc_point_in_cube.restype = C.c_int
    L.orc_ray_tri_intersect.argtypes = [F3, F3, F3, F3, F3]
    L.orc_ray_tri_intersect.restype = C.c_int
    L.orc_populate_cubes.argtypes = [_fp, _up, C.c_uint32, C.c_uint32, _up, _up, C.c_uint32]
    L.orc_populate_cubes.restype = C.c_uint32
    L.orc_populate_aabb.argtypes = [_fp, _up, C.c_uint32, _up, _fp, _fp, C.c_float, _up, _up,
                                    C.c_uint32]
    L.orc_populate_aabb.restype = C.c_uint32
    L.orc_populate_aabb_plane.argtypes = L.orc_populate_aabb.argtypes
    L.orc_populate_aabb_plane.restype = C.c_uint32
    L.orc_tri_coll_detect.argtypes = [P(WorldC), F3, F3, C.c_uint32]
    L.orc_tri_coll_detect.restype = Coll
    L.orc_mt_coll_detect.argtypes = [P(WorldC), F3, F3, C.c_uint32]
    L.orc_mt_coll_detect.restype = Coll
    L.orc_cube_coll_detect.argtypes = [P(WorldC), F3, F3, C.c_uint32, C.c_uint32]
    L.orc_cube_coll_detect.restype = Coll
    L.orc_reflect_pos.argtypes = [P(WorldC), F3, F3, F3, C.c_uint32, C.c_uint32]
    L.orc_reflect_pos.restype = F3
    L.orc_coll_detect_grid.argtypes = [P(WorldC), F3, F3, C.c_float, _u8p, _u16p, P(Hits)]
    L.orc_coll_detect_grid.restype = F3
    L.orc_coll_detect_brute.argtypes = [P(WorldC), F3, F3, C.c_float, _u8p, _u16p, P(Hits)]
    L.orc_coll_detect_brute.restype = F3
    L.orc_integrate.argtypes = [P(WorldC), P(StateC), C.c_uint32, P(StepArgsC)]
    L.orc_find_spins.argtypes = [P(WorldC), _fp, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                 C.c_uint32, C.c_uint32, _u8p, _u16p, _up]
    L.orc_find_spins.restype = C.c_uint32
    L.orc_mr_signal.argtypes = [C.c_uint32, _fp, _fp, _u8p, P(C.c_double)]
    L.orc_moment_signal.argtypes = [C.c_uint32, _fp, _fp, _fp, C.c_float, _u8p, C.c_uint32, _fp,
                                    C.c_float, P(C.c_double)]
    L.orc_parse_mesh.argtypes = [C.c_char_p, C.c_float, P(MeshC)]
    L.orc_parse_mesh.restype = C.c_int
    L.orc_free_mesh.argtypes = [P(MeshC)]
    L.orc_parse_grads.argtypes = [C.c_char_p, _fp, C.c_int]
    L.orc_parse_grads.restype = C.c_int
    L.orc_pgse_schedule.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                    P(C.c_int), C.c_int, P(C.c_int)]
    L.orc_pgse_schedule.restype = C.c_int


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def f3(v) -> F3:
    return F3(float(v[0]), float(v[1]), float(v[2]))


# --------------------------------------------------------------------------- #
# RNG
# --------------------------------------------------------------------------- #
def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32).copy()
    k = np.asarray(key, dtype=np.uint32).copy()
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def mwc_next(seed):
    s = np.asarray(seed, dtype=np.uint32).copy()
    r = lib().orc_mwc_next(_p(s, C.c_uint32))
    return int(r), s


def randn(seed):
    s = np.asarray(seed, dtype=np.uint32).copy()
    v = np.zeros(4, np.float32)
    base = v.ctypes.data
    ptr = [C.cast(base + 4 * k, _fp) for k in range(4)]
    lib().orc_randn(_p(s, C.c_uint32), *ptr)
    return v, s


def u32_to_unit(x):
    return lib().orc_u32_to_unit(int(x))


def philox_step_normals(key, spin, step):
    k = np.asarray(key, np.uint32).copy()
    n = np.zeros(3, np.float32)
    lib().orc_philox_step_normals(_p(k, C.c_uint32), spin, step, _p(n, C.c_float))
    return n


def philox_step_uniform(key, spin, step):
    k = np.asarray(key, np.uint32).copy()
    return lib().orc_philox_step_uniform(_p(k, C.c_uint32), spin, step)


# --------------------------------------------------------------------------- #
# Geometry
# --------------------------------------------------------------------------- #
def tri_helpers(vertices, trgls):
    V = np.ascontiguousarray(vertices, np.float32).ravel()
    T = np.ascontiguousarray(trgls, np.uint32).ravel()
    n = T.size // 3
    H = np.zeros(n * 12, np.float32)
    lib().orc_tri_helpers(_p(V, C.c_float), _p(T, C.c_uint32), n, _p(H, C.c_float))
    return H.reshape(n, 12)


def tri_mt_records(vertices, trgls):
    V = np.ascontiguousarray(vertices, np.float32).ravel()
    T = np.ascontiguousarray(trgls, np.uint32).ravel()
    n = T.size // 3
    R = np.zeros(n * 16, np.float32)
    lib().orc_tri_mt_records(_p(V, C.c_float), _p(T, C.c_uint32), n, _p(R, C.c_float))
    return R.reshape(n, 16)


def populate_cubes(vertices, trgls, n_cubes=50):
    """Reference grid: returns (count[n^3], dense list[n^3, maxc], maxc)."""
    V = np.ascontiguousarray(vertices, np.float32).ravel()
    T = np.ascontiguousarray(trgls, np.uint32).ravel()
    n = T.size // 3
    cnt = np.zeros(n_cubes ** 3, np.uint32)
    maxc = lib().orc_populate_cubes(_p(V, C.c_float), _p(T, C.c_uint32), n, n_cubes,
                                    _p(cnt, C.c_uint32), None, 0)
    lst = np.zeros(max(1, n_cubes ** 3 * maxc), np.uint32)
    lib().orc_populate_cubes(_p(V, C.c_float), _p(T, C.c_uint32), n, n_cubes, _p(cnt, C.c_uint32),
                             _p(lst, C.c_uint32), maxc)
    return cnt, lst[: n_cubes ** 3 * maxc].reshape(n_cubes ** 3, maxc), maxc


def populate_aabb(vertices, trgls, dims, lo, cell, eps, plane=False):
    V = np.ascontiguousarray(vertices, np.float32).ravel()
    T = np.ascontiguousarray(trgls, np.uint32).ravel()
    n = T.size // 3
    d = np.asarray(dims, np.uint32).copy()
    l = np.asarray(lo, np.float32).copy()
    c = np.asarray(cell, np.float32).copy()
    nc = int(d[0]) * int(d[1]) * int(d[2])
    cnt = np.zeros(nc, np.uint32)
    args = [_p(V, C.c_float), _p(T, C.c_uint32), n, _p(d, C.c_uint32), _p(l, C.c_float),
            _p(c, C.c_float), C.c_float(eps)]
    fn = lib().orc_populate_aabb_plane if plane else lib().orc_populate_aabb
    maxc = fn(*args, _p(cnt, C.c_uint32), None, 0)
    lst = np.zeros(max(1, nc * maxc), np.uint32)
    fn(*args, _p(cnt, C.c_uint32), _p(lst, C.c_uint32), maxc)
    return cnt, lst[: nc * maxc].reshape(nc, maxc), maxc


def ray_tri_intersect(r1, r2, t1, t2, t3):
    return bool(lib().orc_ray_tri_intersect(f3(r1), f3(r2), f3(t1), f3(t2), f3(t3)))


class World:
    """Owns the numpy arrays behind an orc_world."""

    def __init__(self, vertices, trgls, tri_info, *, grid="reference", n_cubes=50, dims=None,
                 lo=None, cell=None, eps=None, std_dev=(0, 0, 0, 0), t2=(np.inf,) * 4, dt=0.005,
                 permeability=0.0, reflection_type=1, spin_radius=0.005, boundary=0, test=0,
                 traversal=0, transitions=0, fiber_has_myelin=None, grid_arrays=None,
                 voxel_lo=(0.0, 0.0, 0.0), voxel_hi=(0.0, 0.0, 0.0)):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.trgls = np.ascontiguousarray(trgls, np.uint32).reshape(-1, 3)
        self.tri_info = np.ascontiguousarray(tri_info, np.uint32).reshape(-1, 3)
        self.helpers = tri_helpers(self.vertices, self.trgls)
        self.tri_mt = tri_mt_records(self.vertices, self.trgls)
        n_fibers = int(self.tri_info[:, 0].max()) + 1 if len(self.tri_info) else 0
        if fiber_has_myelin is None:
            fm = np.zeros(max(1, n_fibers), np.uint8)
            for f, m in self.tri_info[:, :2]:
                if m == 1:
                    fm[f] = 1
        else:
            fm = np.ascontiguousarray(fiber_has_myelin, np.uint8)
        self.fiber_has_myelin = fm
        if grid == "reference":
            L = np.float32(2.0) / np.float32(n_cubes)
            self.dims = (n_cubes,) * 3
            self.lo = (-1.0,) * 3
            self.cell = (float(L),) * 3
            self.ref_cell_math = 1
            if grid_arrays is None:
                grid_arrays = populate_cubes(self.vertices, self.trgls, n_cubes)
        else:  # "aabb", or "aabb_plane" (the plane must pass through the cell too)
            self.dims = tuple(int(d) for d in dims)
            self.lo = tuple(float(np.float32(x)) for x in lo)
            self.cell = tuple(float(np.float32(x)) for x in cell)
            self.ref_cell_math = 0
            if grid_arrays is None:
                grid_arrays = populate_aabb(self.vertices, self.trgls, self.dims, self.lo,
                                            self.cell, eps, plane=grid == "aabb_plane")
        self.cell_count, self.cell_list, self.maxc = grid_arrays
        self.cell_count = np.ascontiguousarray(self.cell_count, np.uint32)
        self.cell_list = np.ascontiguousarray(self.cell_list, np.uint32)
        self.inv_cell = tuple(float(np.float32(1.0) / np.float32(c)) for c in self.cell)
        t2a = np.asarray(t2, np.float32)
        self.params = dict(std_dev=np.asarray(std_dev, np.float32), t2=t2a,
                           decay=(np.float32(dt) / t2a).astype(np.float32), dt=np.float32(dt),
                           permeability=np.float32(permeability), reflection_type=reflection_type,
                           spin_radius=np.float32(spin_radius), boundary=boundary, test=test,
                           traversal=traversal, transitions=transitions)
        self.n_fibers = n_fibers
        self.voxel_lo = tuple(float(np.float32(v)) for v in voxel_lo)
        self.voxel_hi = tuple(float(np.float32(v)) for v in voxel_hi)
        self._build()

    def _build(self):
        w = WorldC()
        w.n_tri = len(self.trgls)
        w.helpers = _p(self.helpers, C.c_float)
        w.vertices = _p(self.vertices, C.c_float)
        w.tri_info = _p(self.tri_info, C.c_uint32)
        w.trgls = _p(self.trgls, C.c_uint32)
        w.tri_mt = _p(self.tri_mt, C.c_float)
        w.dims = (C.c_uint32 * 3)(*self.dims)
        w.lo = (C.c_float * 3)(*self.lo)
        w.cell = (C.c_float * 3)(*self.cell)
        w.inv_cell = (C.c_float * 3)(*self.inv_cell)
        w.ref_cell_math = self.ref_cell_math
        w.max_per_cell = int(self.maxc)
        w.cell_count = _p(self.cell_count, C.c_uint32)
        w.cell_list = _p(self.cell_list, C.c_uint32)
        p = self.params
        w.std_dev = (C.c_float * 4)(*[float(x) for x in p["std_dev"]])
        w.t2 = (C.c_float * 4)(*[float(x) for x in p["t2"]])
        w.decay = (C.c_float * 4)(*[float(x) for x in p["decay"]])
        w.dt = float(p["dt"])
        w.permeability = float(p["permeability"])
        w.reflection_type = int(p["reflection_type"])
        w.spin_radius = float(p["spin_radius"])
        w.boundary = int(p["boundary"])
        w.test = int(p["test"])
        w.traversal = int(p["traversal"])
        w.transitions = int(p["transitions"])
        w.fiber_has_myelin = _p(self.fiber_has_myelin, C.c_uint8)
        w.n_fibers = self.n_fibers
        w.wrap_lo = (C.c_float * 3)(*self.voxel_lo)
        w.wrap_hi = (C.c_float * 3)(*self.voxel_hi)
        self.c = w

    def set(self, **kw):
        for k, v in kw.items():
            self.params[k] = v
        if "dt" in kw or "t2" in kw:
            self.params["decay"] = (np.float32(self.params["dt"]) /
                                    np.asarray(self.params["t2"], np.float32)).astype(np.float32)
        self._build()
        return self

    # -- single-call helpers -------------------------------------------------
    def tri_coll(self, o, p, tri):
        return lib().orc_tri_coll_detect(C.byref(self.c), f3(o), f3(p), tri)

    def mt_coll(self, o, p, tri):
        return lib().orc_mt_coll_detect(C.byref(self.c), f3(o), f3(p), tri)

    def reflect(self, s, t, c, tri, typ):
        r = lib().orc_reflect_pos(C.byref(self.c), f3(s), f3(t), f3(c), tri, typ)
        return np.array([r.x, r.y, r.z], np.float32)

    def cube_pos(self, p):
        c = np.zeros(3, np.uint32)
        lib().orc_cube_pos(C.byref(self.c), f3(p), _p(c, C.c_uint32))
        return c

    def cube_hash(self, c):
        cc = np.asarray(c, np.uint32).copy()
        return lib().orc_cube_hash(C.byref(self.c), _p(cc, C.c_uint32))

    def coll_detect(self, o, p, u, comp, fiber, brute=False):
        c8 = C.c_uint8(comp)
        f16 = C.c_uint16(fiber)
        h = Hits()
        fn = lib().orc_coll_detect_brute if brute else lib().orc_coll_detect_grid
        r = fn(C.byref(self.c), f3(o), f3(p), C.c_float(u), C.byref(c8), C.byref(f16), C.byref(h))
        return (np.array([r.x, r.y, r.z], np.float32), c8.value, f16.value,
                (h.first, h.last, h.ncoll))

    def find_spins(self, pos, xmax=1.0, ymax=1.0, zmax=1.0, max_tries=20, plus_x=False,
                   ray_axis=None):
        """ray_axis 0/1/2 casts along +x/+y/+z (plus_x=True is ray_axis 0);
        neither: the reference's dominant-axis ray."""
        if ray_axis is not None:
            plus_x = int(ray_axis) + 1
        P = np.ascontiguousarray(pos, np.float32).ravel()
        n = P.size // 3
        comp = np.zeros(n, np.uint8)
        fib = np.zeros(n, np.uint16)
        tries = np.zeros(n, np.uint32)
        amb = lib().orc_find_spins(C.byref(self.c), _p(P, C.c_float), n, xmax, ymax, zmax,
                                   max_tries, int(plus_x), _p(comp, C.c_uint8), _p(fib, C.c_uint16),
                                   _p(tries, C.c_uint32))
        return comp, fib, tries, amb


class State:
    def __init__(self, pos, comp=None, fiber=None, mag=None, phase=True, seed=None, moment=False,
                 wrap=True):
        self.pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3).copy()
        n = len(self.pos)
        self.n = n
        self.mag = np.ones(n, np.float32) if mag is None else np.asarray(mag, np.float32).copy()
        self.phase = np.zeros(n, np.float32) if phase is True else (
            None if phase is None or phase is False else np.asarray(phase, np.float32).copy())
        self.comp = np.zeros(n, np.uint8) if comp is None else np.asarray(comp, np.uint8).copy()
        self.fiber = (np.full(n, FIBER_NONE, np.uint16) if fiber is None
                      else np.asarray(fiber, np.uint16).copy())
        self.seed = None if seed is None else np.ascontiguousarray(seed, np.uint32).reshape(-1, 2).copy()
        # periodic voxel crossings (wrap=False: the phase of the wrapped position)
        self.wrap = np.zeros((n, 3), np.int32) if wrap else None
        if moment:
            self.disp = np.zeros((n, 3), np.float32)
            self.moment = np.zeros((n, 3), np.float32)
            self.moment_c = np.zeros((n, 3), np.float32)
        else:
            self.disp = self.moment = self.moment_c = None

    def c(self):
        s = StateC()
        s.n = self.n
        s.pos = _p(self.pos, C.c_float)
        s.mag = _p(self.mag, C.c_float)
        s.phase = _p(self.phase, C.c_float)
        s.comp = _p(self.comp, C.c_uint8)
        s.fiber = _p(self.fiber, C.c_uint16)
        s.seed = _p(self.seed, C.c_uint32)
        s.disp = _p(self.disp, C.c_float)
        s.moment = _p(self.moment, C.c_float)
        s.moment_c = _p(self.moment_c, C.c_float)
        s.wrap = _p(self.wrap, C.c_int32)
        return s


def integrate(world: World, state: State, iters, *, rng=RNG_INJECT, inject=None, key=(0, 0),
              step0=0, grad=(0, 0, 0), phase_constant=0.0, weight=0.0, trace=False, n_threads=0,
              stats=None):
    a = StepArgsC()
    a.rng = rng
    a.philox_key = (C.c_uint32 * 2)(*[int(k) for k in key])
    a.step0 = step0
    inj = None
    if inject is not None:
        inj = np.ascontiguousarray(inject, np.float32)
        assert inj.shape == (iters, state.n, 4)
        a.inject = _p(inj, C.c_float)
    a.grad = f3(np.asarray(grad, np.float32))
    a.phase_constant = float(np.float32(phase_constant))
    a.weight = float(weight)
    tr = None
    if trace:
        tr = np.zeros((iters, state.n, 5), np.uint32)
        a.trace = _p(tr, C.c_uint32)
    a.n_threads = n_threads
    if stats is not None:
        assert stats.dtype == np.uint64 and stats.size >= 3
        a.stats = stats.ctypes.data_as(C.POINTER(C.c_uint64))
    sc = state.c()
    lib().orc_integrate(C.byref(world.c), C.byref(sc), iters, C.byref(a))
    return tr


def mr_signal(mag, phase, comp):
    m = np.ascontiguousarray(mag, np.float32)
    p = np.ascontiguousarray(phase, np.float32)
    c = np.ascontiguousarray(comp, np.uint8)
    out = (C.c_double * 5)()
    lib().orc_mr_signal(len(m), _p(m, C.c_float), _p(p, C.c_float), _p(c, C.c_uint8), out)
    return np.array(out[:])


def moment_signal(mag, moment, comp, grads_scaled, phase_constant, x0=None, wsum=0.0):
    m = np.ascontiguousarray(mag, np.float32)
    M = np.ascontiguousarray(moment, np.float32)
    c = np.ascontiguousarray(comp, np.uint8)
    g = np.ascontiguousarray(grads_scaled, np.float32).reshape(-1, 3)
    x = None if x0 is None else np.ascontiguousarray(x0, np.float32)
    out = np.zeros(len(g) * 5, np.float64)
    lib().orc_moment_signal(len(m), _p(m, C.c_float), _p(M, C.c_float), _p(x, C.c_float),
                            float(wsum), _p(c, C.c_uint8), len(g), _p(g, C.c_float),
                            float(np.float32(phase_constant)), _p(out, C.c_double))
    return out.reshape(len(g), 5)


def parse_mesh(path, space_scale):
    m = MeshC()
    rc = lib().orc_parse_mesh(str(path).encode(), float(space_scale), C.byref(m))
    res = dict(rc=rc, n_vert=m.n_vert, n_tri=m.n_tri, n_axon=m.n_axon, n_myelin=m.n_myelin,
               n_glia=m.n_glia, n_fibers=m.n_fibers, max_tri_on_surface=m.max_tri_on_surface,
               xmax=m.xmax, ymax=m.ymax, zmax=m.zmax)
    if m.vertices:
        res["vertices"] = np.ctypeslib.as_array(m.vertices, (m.n_vert * 3,)).reshape(-1, 3).copy() \
            if m.n_vert else np.zeros((0, 3), np.float32)
        res["trgls"] = np.ctypeslib.as_array(m.trgls, (m.n_tri * 3,)).reshape(-1, 3).copy() \
            if m.n_tri else np.zeros((0, 3), np.uint32)
        res["tri_info"] = np.ctypeslib.as_array(m.tri_info, (m.n_tri * 3,)).reshape(-1, 3).copy() \
            if m.n_tri else np.zeros((0, 3), np.uint32)
    lib().orc_free_mesh(C.byref(m))
    return res


def parse_grads(path, cap=4096):
    rows = np.zeros(cap * 6, np.float32)
    n = lib().orc_parse_grads(str(path).encode(), _p(rows, C.c_float), cap)
    return rows[: max(n, 0) * 6].reshape(-1, 6)


def pgse_schedule(ldelta, bdelta, readout, dt, steps_per_update, cap=1 << 22):
    sign = np.zeros(cap, np.int32)
    lbr = np.zeros(4, np.int32)
    n = lib().orc_pgse_schedule(ldelta, bdelta, readout, dt, steps_per_update,
                                _p(sign, C.c_int), cap, _p(lbr, C.c_int))
    return sign[:n].copy(), tuple(int(x) for x in lbr)

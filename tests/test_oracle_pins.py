"""Pins of the oracle (oracle/dsim_oracle.c) against known answers, closed
forms, invariants and exact-rational brute force -- never against a retyped
copy of its own formulas."""
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle as O

# --------------------------------------------------------------------------- #
# RNG
# --------------------------------------------------------------------------- #


def test_philox_random123_known_answers():
    # kat_vectors of the Random123 distribution, philox4x32 with 10 rounds
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0])) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_mwc_recurrence_closed_form():
    # Marsaglia MWC, spinKernelCpu.cpp:41-45, evaluated with Python integers
    s0, s1 = 123456789, 987654321
    seed = [s0, s1]
    for _ in range(50):
        s0 = (36969 * (s0 & 65535) + (s0 >> 16)) & 0xFFFFFFFF
        s1 = (18000 * (s1 & 65535) + (s1 >> 16)) & 0xFFFFFFFF
        want = ((s0 << 16) + s1) & 0xFFFFFFFF
        got, seed = O.mwc_next(seed)
        assert got == want and list(seed) == [s0, s1]


def test_randn_box_muller_closed_form():
    seed = [42, 4242]
    n, _ = O.randn(seed)
    # recompute the four uniforms and the transform in float64
    s = list(seed)
    us = []
    for _ in range(4):
        r, s = O.mwc_next(s)
        us.append((np.float32(r) + np.float32(1.0)) / np.float32(4294967296.0))
    u1, u2, u3, u4 = [float(u) for u in us]
    tp = float(np.float32(6.28318530717959))
    want = [np.sqrt(-2 * np.log(u1)) * np.cos(tp * u2), np.sqrt(-2 * np.log(u1)) * np.sin(tp * u2),
            np.sqrt(-2 * np.log(u3)) * np.cos(tp * u4), u4]
    np.testing.assert_allclose(n, want, rtol=2e-6, atol=2e-6)
    assert n[3] == us[3]  # the fourth uniform is returned untransformed


def test_unit_mapping_open_interval():
    assert O.u32_to_unit(0) == 2.0 ** -24
    assert O.u32_to_unit(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    assert 0.0 < O.u32_to_unit(0x80000000) < 1.0


def test_philox_normals_moments():
    key = (7, 11)
    n = np.array([O.philox_step_normals(key, i, 3) for i in range(20000)])
    assert abs(n.mean()) < 0.03
    np.testing.assert_allclose(n.std(axis=0), 1.0, atol=0.03)
    assert abs(np.corrcoef(n[:, 0], n[:, 1])[0, 1]) < 0.03


def test_philox_uniform_is_the_low_bits_of_the_normals_block():
    key, spin, step = (7, 11), 5, (1 << 33) + 9
    o = O.philox([spin, step & 0xFFFFFFFF, step >> 32, 0], key)
    m = (o[0] & 511) | ((o[1] & 511) << 9) | ((o[2] & 31) << 18)
    assert O.philox_step_uniform(key, spin, step) == (m + 0.5) * 2.0 ** -23
    # uniform on (0,1) and uncorrelated with the normals drawn from the same block
    from scipy import stats
    u = np.array([O.philox_step_uniform(key, i, 3) for i in range(20000)])
    n = np.array([O.philox_step_normals(key, i, 3) for i in range(20000)])
    assert 0.0 < u.min() and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    for k in range(3):
        assert abs(np.corrcoef(u, n[:, k])[0, 1]) < 0.03
        assert abs(np.corrcoef(u, np.abs(n[:, k]))[0, 1]) < 0.03


# --------------------------------------------------------------------------- #
# Triangle tests
# --------------------------------------------------------------------------- #


def _world(vertices, trgls, tri_info=None, **kw):
    vertices = np.asarray(vertices, np.float32)
    trgls = np.asarray(trgls, np.uint32).reshape(-1, 3)
    if tri_info is None:
        tri_info = np.zeros_like(trgls)
        tri_info[:, 2] = trgls[:, 0]
    return O.World(vertices, trgls, tri_info, **kw)


def test_tri_coll_collisions01_known_answer():
    # dSim3/matlab_tests/run_tests/collisions01.m:6-12
    o = (-0.00854051, -0.418389, -0.635162)
    p = (-0.00918993, -0.419038, -0.644214)
    v = [(0.0357, -0.4258, -0.7714), (-0.0308, -0.4167, -0.5751), (-0.0036, -0.3571, -0.5949)]
    w = _world(v, [0, 1, 2], grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1), cell=(2, 2, 2), eps=0)
    # the script's normal (unnormalised cross product) matches the helper normal direction
    n_script = np.array([-0.0118797, 0.00402266, -0.00421092])
    np.testing.assert_allclose(w.helpers[0, :3], n_script / np.linalg.norm(n_script), atol=2e-5)
    r = w.tri_coll(o, p, 0)
    assert r.type == 1 and r.idx == 0
    np.testing.assert_allclose([r.pt.x, r.pt.y, r.pt.z], (-0.00881566, -0.418664, -0.638997),
                               atol=2e-6)


def _exact_seg_tri(o, p, a, b, c):
    """Exact rational segment/triangle intersection: (hit, point, s, t)."""
    o, p, a, b, c = [[Fr(float(x)) for x in v] for v in (o, p, a, b, c)]
    sub = lambda x, y: [x[i] - y[i] for i in range(3)]
    dot = lambda x, y: sum(x[i] * y[i] for i in range(3))
    cross = lambda x, y: [x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]]
    u, v = sub(c, a), sub(b, a)
    n = cross(v, u)
    den = dot(n, sub(p, o))
    if den == 0:
        return False, None, None, None
    r = dot(n, sub(a, o)) / den
    if not (0 < r < 1):
        return False, None, r, None
    q = [o[i] + r * (p[i] - o[i]) for i in range(3)]
    w = sub(q, a)
    uv, uu, vv, wu, wv = dot(u, v), dot(u, u), dot(v, v), dot(w, u), dot(w, v)
    s = (uv * wv - vv * wu) / (uv * uv - uu * vv)
    t = (uv * wu - uu * wv) / (uv * uv - uu * vv)
    return (s >= 0 and t >= 0 and s + t <= 1), q, s, t


def test_tri_coll_exact_rational_brute_force():
    rng = np.random.default_rng(5)
    checked = 0
    for trial in range(400):
        tri = rng.uniform(-0.5, 0.5, (3, 3)).astype(np.float32)
        o = rng.uniform(-0.6, 0.6, 3).astype(np.float32)
        p = (o + rng.normal(0, 0.4, 3)).astype(np.float32)
        w = _world(tri, [0, 1, 2], grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1), cell=(2, 2, 2), eps=0)
        hit, q, s, t = _exact_seg_tri(o, p, *tri)
        # skip near-boundary cases where float rounding may legitimately flip
        if s is not None and t is not None:
            margin = min(abs(s), abs(t), abs(1 - s - t))
            if margin < 1e-4:
                continue
        elif s is not None and min(abs(s), abs(1 - s)) < 1e-4:  # s holds r here
            continue
        r = w.tri_coll(o, p, 0)
        assert (r.type > 0) == hit, trial
        if hit:
            np.testing.assert_allclose([r.pt.x, r.pt.y, r.pt.z], [float(x) for x in q], atol=1e-5)
        checked += 1
    assert checked > 300


def test_tri_coll_edge_and_vertex_classes():
    # right triangle in z=0: (0,0,0),(1,0,0),(0,1,0); vertical segments through (a,b)
    w = _world([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [0, 1, 2], grid="aabb", dims=(1, 1, 1),
               lo=(-1, -1, -1), cell=(2, 2, 2), eps=0)
    cases = {(0.25, 0.25): 1, (0.0, 0.5): 2, (0.5, 0.0): 2, (0.5, 0.5): 2, (0.0, 0.0): 3,
             (1.0, 0.0): 3, (0.0, 1.0): 3, (0.75, 0.5): 0, (-0.25, 0.25): 0}
    for (a, b), typ in cases.items():
        r = w.tri_coll((a, b, 0.5), (a, b, -0.5), 0)
        assert r.type == typ, (a, b, r.type)
    # segment ending exactly on the plane: r == 1 is not a hit (strict 0<r<1)
    assert w.tri_coll((0.25, 0.25, 0.5), (0.25, 0.25, 0.0), 0).type == 0


def test_mt_agrees_with_exact_brute_force():
    rng = np.random.default_rng(11)
    agree = 0
    for trial in range(300):
        tri = rng.uniform(-0.5, 0.5, (3, 3)).astype(np.float32)
        o = rng.uniform(-0.6, 0.6, 3).astype(np.float32)
        p = (o + rng.normal(0, 0.4, 3)).astype(np.float32)
        w = _world(tri, [0, 1, 2], grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1), cell=(2, 2, 2),
                   eps=0, test=1)
        hit, q, s, t = _exact_seg_tri(o, p, *tri)
        if s is not None and t is not None and min(abs(s), abs(t), abs(1 - s - t)) < 1e-4:
            continue
        if s is not None and t is None and min(abs(s), abs(1 - s)) < 1e-4:
            continue
        r = w.mt_coll(o, p, 0)
        assert (r.type > 0) == hit, trial
        if hit:
            np.testing.assert_allclose([r.pt.x, r.pt.y, r.pt.z], [float(x) for x in q], atol=1e-5)
        agree += 1
    assert agree > 200


def test_mt_near_grazing_agrees_with_exact_brute_force():
    # segments within 3e-5..1e-3 rad of the triangle's plane, crossing it
    # inside or just outside the triangle (end points at least 3e-7 from the
    # plane, beyond the rounding of the plane distances): the FP64 re-test must
    # decide them as exact rational arithmetic does on the stored triangle (v0
    # and the float edges e1, e2; at such angles their rounding moves the
    # crossing visibly) and place the hit to float accuracy
    rng = np.random.default_rng(23)
    agree = hits = 0
    for trial in range(400):
        tri = rng.uniform(-0.5, 0.5, (3, 3)).astype(np.float32)
        a, b, c = tri.astype(np.float64)
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        s_, t_ = rng.uniform(-0.3, 1.0, 2)
        x = a + s_ * (b - a) + t_ * (c - a)
        w_ = np.cross(n, rng.normal(size=3))
        w_ /= np.linalg.norm(w_)
        th = 10.0 ** rng.uniform(-4.5, -3)
        d = np.cos(th) * w_ + np.sin(th) * n * rng.choice([-1, 1])
        L, f = rng.uniform(0.05, 0.5), rng.uniform(0.2, 0.8)
        o = (x - f * L * d).astype(np.float32)
        p = (x + (1 - f) * L * d).astype(np.float32)
        stored = [tri[0].astype(np.float64) + (tri[k] - tri[0]).astype(np.float64) for k in (1, 2)]
        hit, q, s, t = _exact_seg_tri(o, p, tri[0], *stored)
        if s is not None and t is not None and min(abs(s), abs(t), abs(1 - s - t)) < 1e-6:
            continue
        if s is not None and t is None and min(abs(s), abs(1 - s)) < 1e-6:
            continue
        w = _world(tri, [0, 1, 2], grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1), cell=(2, 2, 2),
                   eps=0, test=1)
        r = w.mt_coll(o, p, 0)
        assert (r.type > 0) == hit, (trial, th)
        if hit:
            np.testing.assert_allclose([r.pt.x, r.pt.y, r.pt.z], [float(v) for v in q], atol=2e-7)
            hits += 1
        agree += 1
    assert agree > 350 and hits > 100


@pytest.mark.parametrize("test", [0, 1])
def test_reflection_householder_invariants(test):
    rng = np.random.default_rng(3)
    for _ in range(100):
        tri = rng.uniform(-0.5, 0.5, (3, 3)).astype(np.float32)
        w = _world(tri, [0, 1, 2], grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1), cell=(2, 2, 2),
                   eps=0, test=test)
        n = w.helpers[0, :3].astype(np.float64)
        c = tri.mean(axis=0)
        t = (c + rng.normal(0, 0.2, 3)).astype(np.float32)
        s = (c + rng.normal(0, 0.2, 3)).astype(np.float32)
        r = w.reflect(s, t, c, 0, 1).astype(np.float64)
        d_in, d_out = t - c, r - c
        assert abs(np.linalg.norm(d_in) - np.linalg.norm(d_out)) < 1e-6       # isometry
        assert abs(n @ d_in + n @ d_out) < 1e-6                              # normal flips
        np.testing.assert_allclose(d_in - (n @ d_in) * n, d_out - (n @ d_out) * n, atol=1e-6)
        r2 = w.reflect(s, r.astype(np.float32), c, 0, 1)                      # involution
        np.testing.assert_allclose(r2, t, atol=1e-6)
        # edge/vertex hits and reflectionType 0 go back to the start
        np.testing.assert_array_equal(w.reflect(s, t, c, 0, 2), s)


# --------------------------------------------------------------------------- #
# Grid
# --------------------------------------------------------------------------- #


def test_cube_pos_closed_form():
    w = _world([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [0, 1, 2])
    L = np.float32(2.0) / np.float32(50)
    for p in [(-1.0, -1.0, -1.0), (0.0, 0.0, 0.0), (0.99, -0.99, 0.5), (1.5, -3, 0.02), (0.04, 0.08, -0.04)]:
        want = [min(max(int(np.floor((np.float32(x) + np.float32(1)) / L)), 0), 49) for x in p]
        assert list(w.cube_pos(p)) == want
    assert w.cube_hash([1, 2, 3]) == 3 * 2500 + 2 * 50 + 1


def _sat_overlap(tri, bmin, bmax):
    """Signed SAT overlap (exact rationals): >0 overlapping, <0 separated."""
    T = [[Fr(float(x)) for x in v] for v in tri]
    c = [(Fr(float(bmin[i])) + Fr(float(bmax[i]))) / 2 for i in range(3)]
    h = [(Fr(float(bmax[i])) - Fr(float(bmin[i]))) / 2 for i in range(3)]
    v = [[T[k][i] - c[i] for i in range(3)] for k in range(3)]
    e = [[v[1][i] - v[0][i] for i in range(3)], [v[2][i] - v[1][i] for i in range(3)],
         [v[0][i] - v[2][i] for i in range(3)]]
    cross = lambda a, b: [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]
    axes = [[1, 0, 0], [0, 1, 0], [0, 0, 1], cross(e[0], e[1])]
    for i in range(3):
        for j in range(3):
            ax = [0, 0, 0]
            ax[i] = 1
            axes.append(cross(ax, e[j]))
    best = None
    for a in axes:
        n2 = sum(x * x for x in a)
        if n2 == 0:
            continue
        pr = [sum(a[i] * v[k][i] for i in range(3)) for k in range(3)]
        rad = sum(h[i] * abs(a[i]) for i in range(3))
        ov = min(rad - min(pr), max(pr) + rad)  # distance inside the slab (neg = separated)
        ovn = float(ov) / float(n2) ** 0.5
        best = ovn if best is None else min(best, ovn)
    return best


def test_populate_cubes_vs_exact_sat(ref_inputs):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.triangles_curved"), 1.0)
    V, T = m["vertices"], m["trgls"]
    n = 10
    cnt, lst, maxc = O.populate_cubes(V, T, n)
    member = set()
    for cell in range(n ** 3):
        for k in range(cnt[cell]):
            member.add((cell, int(lst[cell, k])))
    L = np.float32(2.0) / np.float32(n)
    rng = np.random.default_rng(0)
    tris = rng.choice(len(T), 60, replace=False)
    checked = 0
    for t in tris:
        tri = V[T[t]]
        lo = np.clip(np.floor((tri.min(0) + 1) / L).astype(int) - 1, 0, n - 1)
        hi = np.clip(np.floor((tri.max(0) + 1) / L).astype(int) + 1, 0, n - 1)
        for x in range(lo[0], hi[0] + 1):
            for y in range(lo[1], hi[1] + 1):
                for z in range(lo[2], hi[2] + 1):
                    bmin = [np.float32(-1) + np.float32(q) * L for q in (x, y, z)]
                    bmax = [np.float32(-1) + np.float32(q + 1) * L for q in (x, y, z)]
                    ov = _sat_overlap(tri, bmin, bmax)
                    if abs(ov) < 1e-6:
                        continue  # touching: rounding decides
                    cell = (z * n + y) * n + x
                    assert ((cell, int(t)) in member) == (ov > 0), (t, x, y, z, ov)
                    checked += 1
    assert checked > 100


def test_populate_aabb_is_superset_of_sat(ref_inputs):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.triangles_curved"), 1.0)
    V, T = m["vertices"], m["trgls"]
    dims, lo, cell = (12, 9, 7), (-1.1, -1.1, -1.1), (2.2 / 12, 2.2 / 9, 2.2 / 7)
    cnt, lst, _ = O.populate_aabb(V, T, dims, lo, cell, 1e-6)
    member = set()
    for c in range(int(np.prod(dims))):
        for k in range(cnt[c]):
            member.add((c, int(lst[c, k])))
    for t in range(0, len(T), 7):
        tri = V[T[t]]
        mn, mx = tri.min(0), tri.max(0)
        for z in range(dims[2]):
            for y in range(dims[1]):
                for x in range(dims[0]):
                    b0 = [np.float32(lo[i]) + np.float32(q) * np.float32(cell[i]) for i, q in enumerate((x, y, z))]
                    b1 = [np.float32(lo[i]) + np.float32(q + 1) * np.float32(cell[i]) for i, q in enumerate((x, y, z))]
                    aabb = all(b0[i] <= mx[i] + 1e-6 and mn[i] - 1e-6 <= b1[i] for i in range(3))
                    c = (z * dims[1] + y) * dims[0] + x
                    assert ((c, t) in member) == aabb


# --------------------------------------------------------------------------- #
# Collision walk
# --------------------------------------------------------------------------- #


def _two_triangle_scene(test=0, traversal=0):
    # T1 crosses the segment at t=0.8 and reaches the x-major earlier cell
    # (x 25, y 26); T2 crosses it at t=0.55 inside cell (26, 26) only.
    V = [(0.036, 0.080, 0.0), (0.080, 0.036, 0.0), (0.058, 0.058, 0.02),
         (0.042, 0.044, 0.0), (0.044, 0.042, 0.0), (0.043, 0.043, 0.02)]
    return _world(V, [0, 1, 2, 3, 4, 5], test=test, traversal=traversal,
                  tri_info=np.array([[0, 0, 0], [0, 0, 3]], np.uint32))


def test_first_cell_rule_known_answer():
    o, p = (0.01, 0.01, 0.01), (0.07, 0.07, 0.01)
    w = _two_triangle_scene()
    cnt = w.cell_count
    assert cnt[(25 * 50 + 26) * 50 + 25] >= 1  # T1 listed in the x-major earlier cell
    end, comp, fib, hits = w.coll_detect(o, p, 0.5, 0, 65535)
    # reference: cell (25,26,25) is walked before (26,26,25), so T1 (t=0.8) wins
    assert hits == (0, 0, 1)
    np.testing.assert_allclose(end, (0.046, 0.046, 0.01), atol=2e-6)
    # nearest-hit traversal picks T2 at t=0.55 and reflects back from (0.043,0.043)
    for test in (0, 1):
        wn = _two_triangle_scene(test=test, traversal=1)
        end, comp, fib, hits = wn.coll_detect(o, p, 0.5, 0, 65535)
        assert hits == (1, 1, 1)
        np.testing.assert_allclose(end, (0.016, 0.016, 0.01), atol=2e-6)
        e2, _, _, h2 = wn.coll_detect(o, p, 0.5, 0, 65535, brute=True)
        assert h2 == hits and np.array_equal(e2, end)


def _slab(perm, test=0, traversal=0):
    # two planes x=0.1 (axon wall, fibre 3) and x=0.3 (myelin wall, fibre 3)
    V = []
    for x in (0.1, 0.3):
        V += [(x, -0.5, -0.5), (x, 0.5, -0.5), (x, 0.5, 0.5), (x, -0.5, 0.5)]
    T = [0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7]
    info = np.array([[3, 0, 0], [3, 0, 0], [3, 1, 4], [3, 1, 4]], np.uint32)
    return _world(V, T, tri_info=info, permeability=perm, test=test, traversal=traversal,
                  transitions=test)


@pytest.mark.parametrize("test,traversal", [(0, 0), (1, 1)])
def test_reflect_and_transmit_known_answers(test, traversal):
    o, p = (0.0, 0.05, 0.07), (0.4, 0.05, 0.07)
    end, comp, fib, hits = _slab(0.0, test, traversal).coll_detect(o, p, 0.7, 0, 65535)
    np.testing.assert_allclose(end, (-0.2, 0.05, 0.07), atol=1e-6)   # mirrored at x=0.1
    assert hits[2] == 1 and comp == 0
    # permeability 1: u > 1-(1-0)*1 = 0 always transmits through both walls
    end, comp, fib, hits = _slab(1.0, test, traversal).coll_detect(o, p, 0.7, 0, 65535)
    np.testing.assert_allclose(end, p, atol=1e-7)
    assert hits[2] == 2
    if test == 0:
        # reference table: 0 -(axon wall)-> 3 (fibre 3) -> 0 (fibre none)
        assert comp == 0 and fib == 65535
    else:
        # physical table: 0 -(axon)-> 1 -> myelin present -> 2
        assert comp == 2 and fib == 3
    # one uniform per step: second hit transmits iff u > thr2 (u-interval rule)
    w = _slab(0.5, test, traversal)
    _, comp, fib, hits = w.coll_detect(o, p, 0.7, 0, 65535)   # transmit (0.7>0.5), then reflect (0.7<=0.75)
    assert hits[2] == 2


def test_reference_transition_table():
    # reference table spinKernelCpu.cpp:251-270 driven through single hits
    w = _slab(1.0)
    o, p = (0.0, 0.05, 0.07), (0.2, 0.05, 0.07)  # crosses the axon wall only
    for c0, f0, c1, f1 in [(0, 65535, 3, 3), (1, 3, 2, 3), (2, 3, 1, 3), (3, 3, 0, 65535)]:
        _, comp, fib, _ = w.coll_detect(o, p, 0.7, c0, f0)
        assert (comp, fib) == (c1, f1)
    o, p = (0.2, 0.05, 0.07), (0.4, 0.05, 0.07)  # crosses the myelin wall only
    for c0, f0, c1, f1 in [(0, 65535, 2, 3), (2, 3, 0, 65535), (1, 3, 2, 3)]:
        _, comp, fib, _ = w.coll_detect(o, p, 0.7, c0, f0)
        assert (comp, fib) == (c1, f1)


def test_nearest_mode_grid_independent(ref_inputs):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 3.0)
    rng = np.random.default_rng(9)
    w = O.World(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=(17, 13, 5),
                lo=(-1.3, -1.3, -1.3), cell=(2.6 / 17, 2.6 / 13, 2.6 / 5), eps=1e-5, test=1,
                traversal=1, permeability=0.3)
    for _ in range(400):
        o = rng.uniform(-0.9, 0.9, 3).astype(np.float32)
        p = (o + rng.normal(0, 0.15, 3)).astype(np.float32)
        u = float(rng.uniform())
        a = w.coll_detect(o, p, u, 0, 65535)
        b = w.coll_detect(o, p, u, 0, 65535, brute=True)
        assert a[3] == b[3] and a[1:3] == b[1:3]
        np.testing.assert_array_equal(a[0], b[0])


# --------------------------------------------------------------------------- #
# Whole steps and labelling
# --------------------------------------------------------------------------- #


def _poly_contains(poly, x, y):
    inside = False
    n = len(poly)
    for i in range(n):
        x1, y1 = poly[i]
        x2, y2 = poly[(i + 1) % n]
        if (y1 > y) != (y2 > y):
            xi = x1 + (y - y1) * (x2 - x1) / (y2 - y1)
            if x < xi:
                inside = not inside
    return inside


def _rings(m):
    """xy polygons of the A0 and M0 prisms of a sample mesh (ordered by angle)."""
    out = {}
    for mem in (0, 1):
        tris = np.where(m["tri_info"][:, 1] == mem)[0]
        vids = np.unique(m["trgls"][tris].ravel())
        xy = m["vertices"][vids][:, :2]
        xy = np.unique(np.round(xy, 6), axis=0)
        xy = xy[np.hypot(xy[:, 0], xy[:, 1]) > 1e-6]
        ang = np.arctan2(xy[:, 1], xy[:, 0])
        out[mem] = xy[np.argsort(ang)]
    return out


def test_find_spins_matches_point_in_prism(ref_inputs):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 3.0)
    w = O.World(m["vertices"], m["trgls"], m["tri_info"])
    rings = _rings(m)
    rng = np.random.default_rng(2)
    pos = rng.uniform(-0.95, 0.95, (3000, 3)).astype(np.float32)
    comp, fib, tries, amb = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"])
    assert amb == 0
    n_checked = 0
    for i, (x, y, z) in enumerate(pos):
        ina = _poly_contains(rings[0], x, y)
        inm = _poly_contains(rings[1], x, y)
        # skip points within 1e-3 of a wall
        d = min(np.min(np.hypot(*(rings[k] - (x, y)).T)) for k in (0, 1))
        if d < 1e-3:
            continue
        want = 1 if ina else (2 if inm else 0)
        assert comp[i] == want, (i, x, y, z)
        assert fib[i] == (0 if want else 65535)
        n_checked += 1
    assert n_checked > 2500
    assert set(np.unique(comp)) == {0, 1, 2}


def test_integrate_free_diffusion_variance():
    # no membranes: displacement variance after k steps is k*sigma^2 per axis
    w = _world(np.zeros((3, 3), np.float32) + 5, [0, 1, 2], std_dev=(0.01,) * 4)
    st = O.State(np.zeros((4000, 3)), moment=False)
    O.integrate(w, st, 50, rng=O.RNG_PHILOX, key=(3, 4))
    var = st.pos.var(axis=0)
    np.testing.assert_allclose(var, 50 * 1e-4, rtol=0.06)


def test_integrate_impermeable_keeps_labels(ref_inputs):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 3.0)
    w = O.World(m["vertices"], m["trgls"], m["tri_info"], std_dev=(0.02,) * 4)
    rng = np.random.default_rng(4)
    pos = rng.uniform(-0.8, 0.8, (1500, 3)).astype(np.float32)
    comp, fib, _, _ = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"])
    st = O.State(pos, comp=comp, fiber=fib)
    O.integrate(w, st, 40, rng=O.RNG_PHILOX, key=(1, 2))
    np.testing.assert_array_equal(st.comp, comp)
    comp2, _, _, _ = w.find_spins(st.pos, m["xmax"], m["ymax"], m["zmax"])
    assert np.mean(comp2 == comp) > 0.999  # no spin leaks through a closed wall


def test_t2_forward_euler_closed_form():
    w = _world(np.zeros((3, 3), np.float32) + 5, [0, 1, 2], std_dev=(0.0,) * 4,
               t2=(80.0, 70.0, 7.5, 70.0), dt=0.005)
    st = O.State(np.zeros((4, 3)), comp=[0, 1, 2, 3])
    O.integrate(w, st, 200, rng=O.RNG_PHILOX)
    want = (1 - 0.005 / np.array([80.0, 70.0, 7.5, 70.0])) ** 200
    np.testing.assert_allclose(st.mag, want, rtol=1e-5)


def test_mr_signal_closed_form():
    mag = np.ones(6, np.float32)
    phase = np.array([0, np.pi, 0.3, 0.3, 1.0, -1.0], np.float32)
    comp = np.array([0, 0, 1, 1, 2, 2], np.uint8)
    s = O.mr_signal(mag, phase, comp)
    assert s[1] == pytest.approx(0.0, abs=1e-7)                     # opposite phases cancel
    assert s[2] == pytest.approx(1.0)                               # equal phases
    assert s[3] == pytest.approx(abs(np.cos(np.float32(1.0))))      # +-1 rad
    assert s[4] == 0.0                                              # empty compartment
    tot = abs(np.sum(np.exp(1j * phase.astype(np.float64)))) / 6
    assert s[0] == pytest.approx(tot, rel=1e-12)


def test_moment_signal_equals_phase_signal():
    rng = np.random.default_rng(1)
    mom = rng.normal(0, 1, (500, 3)).astype(np.float32)
    mag = rng.uniform(0.5, 1, 500).astype(np.float32)
    comp = rng.integers(0, 4, 500).astype(np.uint8)
    g = np.array([[0.3, -0.2, 0.1], [0, 0, 0]], np.float32)
    out = O.moment_signal(mag, mom, comp, g, 2.0)
    ph = (mom.astype(np.float64) @ g.T.astype(np.float64)) * 2.0
    for j in range(2):
        want = O.mr_signal(mag, ph[:, j].astype(np.float32), comp)
        np.testing.assert_allclose(out[j], want, rtol=1e-5, atol=1e-6)


# --------------------------------------------------------------------------- #
# Formats and sequencing
# --------------------------------------------------------------------------- #


def test_parse_sample_meshes(ref_inputs):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 60.0)
    assert (m["rc"], m["n_vert"], m["n_tri"], m["n_axon"], m["n_myelin"], m["n_glia"]) == (0, 32, 56, 1, 1, 0)
    np.testing.assert_allclose(m["vertices"][0], np.float32([-1.946, -0.9372, -60]) / np.float32(60))
    assert list(m["trgls"][0]) == [8, 14, 12] and list(m["trgls"][-1]) == [30, 27, 31]
    assert (m["tri_info"][:28, :2] == [0, 0]).all() and (m["tri_info"][28:, :2] == [0, 1]).all()
    assert (m["tri_info"][:, 2] == m["trgls"][:, 0]).all()
    c = O.parse_mesh(os.path.join(ref_inputs, "sim.triangles_curved"), 1.0)
    assert (c["rc"], c["n_vert"], c["n_tri"], c["max_tri_on_surface"]) == (0, 304, 600, 300)
    # the header word Max_triangles_on_membrane opens a ninth M section and the
    # reference rejects the file (spinSystem.cpp:593-596)
    h = O.parse_mesh(os.path.join(ref_inputs, "sim.triangles_8f_1g"), 1.0)
    assert (h["rc"], h["n_axon"], h["n_myelin"], h["n_glia"]) == (-2, 8, 9, 1)
    assert O.parse_mesh(os.path.join(ref_inputs, "does_not_exist"), 1.0)["rc"] == -1


def test_parse_grads(ref_inputs):
    g = O.parse_grads(os.path.join(ref_inputs, "sim.grads"))
    assert g.shape == (13, 6)
    assert list(g[0]) == [10.0, np.float32(12.0669), 1.0, 0, 0, 0]
    assert list(g[5, 3:]) == [80.0, 0.0, 80.0]


def test_pgse_schedule_closed_form():
    sign, (L, B, R, ro) = O.pgse_schedule(10.0, 12.0669, 1.0, 0.005, 10)
    assert (L, B, R, ro) == (200, 241, 20, 1)
    assert len(sign) == 461
    assert (sign[:200] == 1).all() and (sign[200:241] == 0).all()
    assert (sign[241:441] == -1).all() and (sign[441:] == 0).all()
    # delta >= Delta: lobes overlap, no off period between them
    sign, (L, B, R, ro) = O.pgse_schedule(10.0, 5.0, 1.0, 0.5, 1)
    assert (L, B, R) == (20, 10, 2)
    assert (sign[:10] == 1).all() and (sign[10:30] == -1).all() and (sign[30:] == 0).all()
    # R = 0: the B+L branch claims the readout update, so no readout happens
    _, (_, _, _, ro) = O.pgse_schedule(10.0, 40.0, 0.0, 0.5, 1)
    assert ro == 0


def test_find_spins_plus_x_mt_matches_point_in_prism(ref_inputs):
    # the physical mode's labeller: Moller-Trumbore crossings of a +x ray
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 3.0)
    w = O.World(m["vertices"], m["trgls"], m["tri_info"], test=1, traversal=1, transitions=1)
    rings = _rings(m)
    rng = np.random.default_rng(5)
    pos = rng.uniform(-0.95, 0.95, (3000, 3)).astype(np.float32)
    comp, fib, tries, amb = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"], plus_x=True)
    assert amb == 0
    n_checked = 0
    for i, (x, y, z) in enumerate(pos):
        d = min(np.min(np.hypot(*(rings[k] - (x, y)).T)) for k in (0, 1))
        if d < 1e-3:
            continue
        ina = _poly_contains(rings[0], x, y)
        inm = _poly_contains(rings[1], x, y)
        want = 1 if ina else (2 if inm else 0)
        assert comp[i] == want, (i, x, y, z)
        assert fib[i] == (0 if want else 65535)
        n_checked += 1
    assert n_checked > 2500
    assert set(np.unique(comp)) == {0, 1, 2}


def test_find_spins_plus_x_mt_on_a_packing():
    # open 32-gon prisms of a periodic packing (with their images around the
    # voxel): every spin away from a wall gets the fibre whose polygon holds it,
    # and the intra-axonal fraction equals the polygons' area fraction
    from dsim_amd import substrate as S
    rng = np.random.default_rng(21)
    ax = S.make_axons(40, 3.0, 2.0, rng, fraction=0.55, ngon=32)
    V, T, I, s = S.prism_mesh(ax, ngon=32, z_range=(-1.2 * ax.voxel / 2, 1.2 * ax.voxel / 2),
                              halo=0.5)
    w = O.World(V, T, I, test=1, traversal=1, transitions=1, grid="aabb", dims=(1, 1, 1),
                lo=(-2, -2, -2), cell=(4, 4, 4), eps=0.0)
    n = 4000
    pos = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    xmax = float(np.abs(V[:, 0]).max())
    comp, fib, tries, amb = w.find_spins(pos, xmax, 1.0, 1.0, plus_x=True)
    assert amb == 0
    ang = 2 * np.pi * np.arange(32) / 32
    polys = []
    for f, ((cx, cy), R) in enumerate(zip(ax.centers, ax.radii)):
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                x0, y0 = (cx - s + ix * ax.voxel) / s, (cy - s + iy * ax.voxel) / s
                polys.append((f, np.stack([x0 + R / s * np.cos(ang), y0 + R / s * np.sin(ang)], 1)))
    checked = 0
    for i, (x, y, _) in enumerate(pos):
        want_f, near = 65535, False
        for f, P in polys:
            if np.abs(P[:, 0].mean() - x) > 0.5 or np.abs(P[:, 1].mean() - y) > 0.5:
                continue
            # distance to the polygon's edges
            a, b = P, np.roll(P, -1, 0)
            ab = b - a
            t = np.clip(((x - a[:, 0]) * ab[:, 0] + (y - a[:, 1]) * ab[:, 1]) / (ab ** 2).sum(1), 0, 1)
            d = np.hypot(a[:, 0] + t * ab[:, 0] - x, a[:, 1] + t * ab[:, 1] - y).min()
            near |= d < 1e-3
            if _poly_contains(P, x, y):
                want_f = f
        if near:
            continue
        assert fib[i] == want_f, (i, x, y)
        assert comp[i] == (1 if want_f != 65535 else 0)
        checked += 1
    assert checked > 0.95 * n
    frac = np.mean(comp == 1)
    sigma = np.sqrt(ax.intra_fraction * (1 - ax.intra_fraction) / n)
    assert abs(frac - ax.intra_fraction) < 3 * sigma + 0.01, (frac, ax.intra_fraction)


def _shared_diagonal_case():
    """A step of the small configuration's walk (16 capped 32-gon prisms, the
    reference 50^3 grid) whose segment meets a prism facet on the diagonal its
    two coplanar triangles share; the spin's state and draw were taken from a
    1e9-spin-step oracle run with the reference's MWC stream."""
    from dsim_amd import substrate as S
    cfg = S.SMALL
    env = S.build(cfg, cell_um=0.2, axons=S.pack(cfg))
    pos = np.array([[-0.01720589, -0.5001226, -0.9633706]], np.float32)
    q = np.array([[[2.1228313, 0.6740799, -1.0266286, 0.646406]]], np.float32)
    return env, pos, q


def test_reference_collision_loop_never_ends_on_a_shared_facet_diagonal():
    # collDetectRectGrid (spinKernelCpu.cpp:209-277) excludes only the last
    # triangle hit: the softsurfer test reports the hit on one triangle of the
    # facet, the reflected segment hits its coplanar twin at t = 0, and back,
    # without end.  The oracle stops at its 100000-iteration guard.
    env, pos, q = _shared_diagonal_case()
    V, T, I = env["vertices"], env["triangles"], env["tri_info"]
    w = O.World(V, T, I, std_dev=env["std_dev"], t2=env["t2"], dt=0.005, permeability=0.0)
    st = O.State(pos, comp=np.zeros(1, np.uint8), fiber=np.full(1, 65535, np.uint16))
    tr = O.integrate(w, st, 1, rng=O.RNG_INJECT, inject=q, trace=True)
    assert tr[0, 0, 2] == 100000                  # the guard, not a finished step
    assert {int(tr[0, 0, 0]), int(tr[0, 0, 1])} == {1200, 1168}
    P0, P1 = V[T[1200]].astype(np.float64), V[T[1168]].astype(np.float64)
    n0 = np.cross(P0[1] - P0[0], P0[2] - P0[0])
    n1 = np.cross(P1[1] - P1[0], P1[2] - P1[0])
    assert np.linalg.norm(np.cross(n0, n1)) < 1e-9 * np.linalg.norm(n0) * np.linalg.norm(n1)
    assert len(set(T[1200].tolist()) & set(T[1168].tolist())) == 2   # one shared edge
    # the physical mode's strict side test does not count a segment that
    # starts on the plane: the same step ends
    w.set(test=1, traversal=1, transitions=1)
    st = O.State(pos, comp=np.zeros(1, np.uint8), fiber=np.full(1, 65535, np.uint16))
    tr = O.integrate(w, st, 1, rng=O.RNG_INJECT, inject=q, trace=True)
    assert tr[0, 0, 2] < 10

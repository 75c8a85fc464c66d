"""Crossing-bundle substrates on the CPU: the generator's packing and mesh,
the exact-membership grid (host build) against exact-rational separating
axes, its clearance against brute-force distances, the labelling ray axis,
and the oracle's +y labels against the analytic geometry."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import crossing as X
from dsim_amd import lib as D
from dsim_amd import substrate as S

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_pins import _sat_overlap  # noqa: E402


def _small(angle, n=60, r=1.0, seed=3):
    rng = np.random.default_rng(seed)
    cx = X.make_crossing(n, r, angle, rng)
    V, T, I, s = X.crossing_mesh(cx, halo=1.5)
    return cx, V, T, I, s


def test_deposit_strip_non_overlapping():
    rng = np.random.default_rng(0)
    r = S.gamma_radii(400, 3.0, 1.0, rng)
    W = 30.0
    pos, h = X.deposit_strip(r, W, rng)
    assert np.all(pos[:, 1] - r >= 0.0) and np.all(pos[:, 1] + r <= h)
    d = pos[:, None, :] - pos[None, :, :]
    d[..., 0] -= W * np.round(d[..., 0] / W)
    dist = np.hypot(d[..., 0], d[..., 1])
    need = r[:, None] + r[None, :] + 0.02 - 1e-9
    np.fill_diagonal(dist, np.inf)
    assert np.all(dist >= need)
    frac = np.pi * np.sum(r ** 2) / (W * h)
    assert 0.45 < frac < 0.8


@pytest.mark.parametrize("angle", [0, 26.6, 45, 63.4, 90])
def test_lattice_directions(angle):
    (p, q), a = X.lattice_direction(angle)
    assert abs(a - angle) < 0.1
    assert np.degrees(np.arctan2(p, q)) == pytest.approx(a)


def test_label_axis_choice(ref_inputs):
    # z-aligned prisms (capped sample mesh, open packing) keep +x; a crossing
    # at any angle but 0 moves the rays to +y, across both bundles
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 3.0)
    assert D.Substrate(m["vertices"], m["trgls"], m["tri_info"]).label_axis() == 0
    ax = S.make_axons(30, 3.0, 2.0, np.random.default_rng(1), fraction=0.5, ngon=32)
    V, T, I, s = S.prism_mesh(ax, ngon=32, z_range=(-20, 20), halo=0.5)
    assert D.Substrate(V, T, I).label_axis() == 0
    for angle, want in ((0, 0), (26.6, 1), (45, 1), (63.4, 1), (90, 1)):
        cx, V, T, I, s = _small(angle)
        assert D.Substrate(V, T, I).label_axis() == want, angle


def test_tribox_membership_against_exact_separating_axes():
    # every cell the triangle truly meets is listed; every listed cell is
    # within the margin of the triangle (exact rationals)
    cx, V, T, I, s = _small(45, n=20)
    g = X.grid_params(cx, 1.5, 2.0)
    sub = D.Substrate(V, T, I, grid="tribox", **g)
    start, tri, _ = sub.grid()
    member = set()
    for c in range(len(start) - 1):
        for k in range(start[c], start[c + 1]):
            member.add((c, int(tri[k])))
    dims, lo, cell, eps = g["dims"], np.float32(g["lo"]), np.float32(g["cell"]), g["eps"]
    rng = np.random.default_rng(5)
    checked = 0
    for t in rng.choice(len(T), 25, replace=False):
        P3 = V[T[t]].astype(np.float64)
        i0 = np.clip(np.floor((P3.min(0) - lo) / cell).astype(int) - 1, 0, np.array(dims) - 1)
        i1 = np.clip(np.floor((P3.max(0) - lo) / cell).astype(int) + 1, 0, np.array(dims) - 1)
        for z in range(i0[2], i1[2] + 1):
            for y in range(i0[1], i1[1] + 1):
                for x in range(i0[0], i1[0] + 1):
                    idx = np.array([x, y, z], np.float64)
                    b0 = lo.astype(np.float64) + idx * cell.astype(np.float64)
                    b1 = lo.astype(np.float64) + (idx + 1) * cell.astype(np.float64)
                    ov = _sat_overlap(P3, b0, b1)
                    c = (z * dims[1] + y) * dims[0] + x
                    listed = (c, int(t)) in member
                    if ov > 0:
                        assert listed, (t, x, y, z, ov)
                    if listed:
                        assert ov > -eps * 1.0001, (t, x, y, z, ov)
                    checked += 1
    assert checked > 300


def _point_tri_dist(p, a, b, c):
    """Distance from p to triangle abc (brute force over a fine barycentric
    sample plus the edges; a lower-accuracy upper bound is not used)."""
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    n2 = n @ n
    # projection inside?
    w = p - a
    s_ = (np.cross(w, ac) @ n) / n2
    t_ = (np.cross(ab, w) @ n) / n2
    if s_ >= 0 and t_ >= 0 and s_ + t_ <= 1:
        return abs(w @ n) / np.sqrt(n2)
    best = np.inf
    for u, v in ((a, b), (b, c), (c, a)):
        e = v - u
        tt = np.clip((p - u) @ e / (e @ e), 0, 1)
        best = min(best, np.linalg.norm(p - (u + tt * e)))
    return best


@pytest.mark.parametrize("rings", [0, 5])
def test_tribox_clearance_is_a_lower_bound(rings):
    # the default (ring distance only) and exact refinement of rings 0-3
    cx, V, T, I, s = _small(63.4, n=20)
    g = X.grid_params(cx, 1.5, 1.5)
    sub = D.Substrate(V, T, I, grid="tribox", clearance_rings=rings, **g)
    cl = sub.clearance()
    dims, lo, cell = g["dims"], np.array(g["lo"]), np.array(g["cell"])
    rng = np.random.default_rng(7)
    P = V[T].astype(np.float64)
    good = 0
    for c in rng.choice(np.where((cl > 0) & (cl < 1e29))[0], 40, replace=False):
        x, y, z = c % dims[0], (c // dims[0]) % dims[1], c // (dims[0] * dims[1])
        b0 = lo + np.array([x, y, z]) * cell
        pt = b0 + rng.uniform(0, 1, 3) * cell
        # only triangles whose box comes near the point need a distance
        near = np.all((P.min(1) <= pt + cl[c] + cell.max()) & (P.max(1) >= pt - cl[c] - cell.max()), 1)
        d = min((_point_tri_dist(pt, *P[k]) for k in np.where(near)[0]), default=np.inf)
        assert d >= cl[c] * (1 - 1e-6), (c, d, cl[c])
        good += 1
    assert good == 40


@pytest.mark.parametrize("angle", [26.6, 90])
def test_oracle_plus_y_labels_match_crossing_geometry(angle):
    cx, V, T, I, s = _small(angle, n=40)
    w = O.World(V, T, I, test=1, traversal=1, transitions=1, grid="aabb", dims=(1, 1, 1),
                lo=(-2, -2, -2), cell=(4, 4, 4), eps=0.0)
    lo, hi = X.voxel_bounds(cx)
    rng = np.random.default_rng(9)
    n = 1500
    pos = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    ymax = float(np.abs(V[:, 1]).max())
    comp, fib, tries, amb = w.find_spins(pos, 1.0, ymax, 1.0, ray_axis=1)
    want, dist = X.inside(cx, pos.astype(np.float64) * s)
    ok = dist > 1e-3 * s
    assert ok.mean() > 0.95
    got = np.where(comp == 1, fib.astype(int), -1)
    np.testing.assert_array_equal(got[ok], want[ok])
    assert 0.2 < np.mean(want >= 0) < 0.8


def test_oracle_plane_grid_is_a_superset_of_exact_overlap():
    # the oracle's own accelerating grid for tilted meshes: every cell a
    # triangle truly meets is listed (exact rationals)
    cx, V, T, I, s = _small(45, n=20)
    dims, lo, cell = (14, 16, 14), (-1.1, -1.2, -1.1), (2.2 / 14, 2.4 / 16, 2.2 / 14)
    cnt, lst, _ = O.populate_aabb(V, T, dims, lo, cell, 1e-6, plane=True)
    cnt_a, _, _ = O.populate_aabb(V, T, dims, lo, cell, 1e-6)
    assert cnt.sum() < 0.5 * cnt_a.sum()  # the plane drops most of a tilted box
    member = set()
    for c in range(int(np.prod(dims))):
        for k in range(cnt[c]):
            member.add((c, int(lst[c, k])))
    rng = np.random.default_rng(4)
    checked = 0
    for t in rng.choice(len(T), 15, replace=False):
        P3 = V[T[t]].astype(np.float64)
        for z in range(dims[2]):
            for y in range(dims[1]):
                for x in range(dims[0]):
                    b0 = [np.float32(lo[i]) + np.float32(q) * np.float32(cell[i]) for i, q in enumerate((x, y, z))]
                    b1 = [np.float32(lo[i]) + np.float32(q + 1) * np.float32(cell[i]) for i, q in enumerate((x, y, z))]
                    if not all(b0[i] <= P3[:, i].max() and P3[:, i].min() <= b1[i] for i in range(3)):
                        continue
                    if _sat_overlap(P3, b0, b1) > 1e-9:
                        assert ((z * dims[1] + y) * dims[0] + x, int(t)) in member
                        checked += 1
    assert checked > 50


def test_oracle_walk_is_grid_independent_on_crossing_mesh():
    # nearest-hit walks over the box grid and the plane grid give the same bits
    cx, V, T, I, s = _small(63.4, n=25)
    lo, hi = X.voxel_bounds(cx)
    sd = np.float32(0.3 / s)
    kw = dict(std_dev=(sd,) * 4, permeability=0.2, boundary=1, test=1, traversal=1,
              transitions=1, voxel_lo=tuple(lo), voxel_hi=tuple(hi), eps=1e-6)
    ext = np.array([cx.E, cx.H, cx.E]) / 2 / s + 1.5 / s
    worlds = [O.World(V, T, I, grid=g, dims=d, lo=tuple(-ext), cell=tuple(2 * ext / np.array(d)), **kw)
              for g, d in (("aabb", (5, 5, 5)), ("aabb_plane", (20, 24, 20)))]
    rng = np.random.default_rng(2)
    pos = rng.uniform(lo, hi, (300, 3)).astype(np.float32)
    comp, fib, _, _ = worlds[0].find_spins(pos, *np.abs(V).max(0), ray_axis=0)
    outs = []
    for w in worlds:
        st = O.State(pos, comp=comp, fiber=fib)
        tr = O.integrate(w, st, 40, rng=O.RNG_PHILOX, key=(5, 6), trace=True)
        outs.append((tr[..., [0, 1, 2, 4]], st.pos, st.comp))
    assert (outs[0][0][..., 2] > 0).sum() > 100
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)

"""Crossing-bundle substrates: two bundles of open N-gon prisms crossing at an
angle, in a periodic voxel.

Layout.  The voxel is E x H x E micrometres (x, y, z), periodic on every
axis.  Bundle A runs along z and fills the slab y in [-H/2, 0); bundle B runs
along d = (p, 0, q)/|(p, q)| and fills the slab y in [0, H/2).  Both axes lie
in the x-z plane, so the bundles never meet, and y is perpendicular to both
(the labelling rays run along +y).

Periodicity.  A straight tube is periodic in the voxel only if its direction
is a lattice direction: (p, 0, q) with integers p, q.  The images of one
tube of bundle B under the voxel's translations are then the parallel lines
offset by multiples of P = E / |(p, q)| along u = (q, 0, -p)/|(p, q)|, so
bundle B is a disk packing in the strip u in [0, P) (periodic), y in [0, H/2)
(walls), and each disk becomes the lines u_i + k P that pass through the
voxel (plus its halo), each meshed as one prism over its chord through the
halo box.  The crossing angle atan(p / q) is drawn from the lattice
directions with p, q <= 2: 0, 26.6, 45, 63.4 and 90 degrees.

Everything returned is normalised by the space scale s = max(E, H) / 2.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.spatial import cKDTree

from . import substrate as S

# lattice directions (p, q) of bundle B and their crossing angles
DIRECTIONS = ((0, 1), (1, 2), (1, 1), (2, 1), (1, 0))


def lattice_direction(angle_deg):
    """The lattice direction (p, q), p, q <= 2, nearest to an angle from z."""
    angs = [np.degrees(np.arctan2(p, q)) for p, q in DIRECTIONS]
    k = int(np.argmin([abs(a - angle_deg) for a in angs]))
    return DIRECTIONS[k], angs[k]


def pack_strip(radii, width, height, rng, gap=0.02, max_iter=20000, tol=1e-9):
    """Non-overlapping disks in a strip periodic in x (period `width`) and
    walled in y (each disk inside [0, height]).  Random placement, then
    pairwise push-apart with the periodic minimum image in x."""
    r = np.asarray(radii, np.float64)
    n = len(r)
    W, h = float(width), float(height)
    rmax = r.max()
    lo_y, hi_y = r + 0.5 * gap, h - r - 0.5 * gap
    if np.any(hi_y < lo_y) or 2 * rmax + gap > W:
        raise ValueError("strip too small for the largest disk")
    pos = np.stack([rng.uniform(0.0, W, n), rng.uniform(lo_y, hi_y)], 1)
    # y never wraps: the box's y period exceeds the strip by two diameters
    box = np.array([W, h + 4 * rmax + 2 * gap])
    for it in range(max_iter):
        tree = cKDTree(pos, boxsize=box)
        pairs = tree.query_pairs(2 * rmax + gap, output_type="ndarray")
        if len(pairs) == 0:
            return pos, it
        i, j = pairs[:, 0], pairs[:, 1]
        d = pos[j] - pos[i]
        d[:, 0] -= W * np.round(d[:, 0] / W)
        dist = np.hypot(d[:, 0], d[:, 1])
        ov = r[i] + r[j] + gap - dist
        m = ov > tol
        if not m.any():
            return pos, it
        i, j, d, dist, ov = i[m], j[m], d[m], dist[m], ov[m]
        u = d / np.maximum(dist, 1e-12)[:, None]
        wi = r[j] ** 2 / (r[i] ** 2 + r[j] ** 2)
        step = (ov * 0.5 + 1e-9)[:, None] * 1.02
        delta = np.zeros_like(pos)
        np.add.at(delta, i, -u * step * 2 * wi[:, None])
        np.add.at(delta, j, u * step * 2 * (1.0 - wi)[:, None])
        pos = pos + delta
        pos[:, 0] = np.mod(pos[:, 0], W)
        pos[:, 1] = np.clip(pos[:, 1], lo_y, hi_y)
    raise RuntimeError(f"strip packing did not converge in {max_iter} iterations")


def deposit_strip(radii, width, rng, gap=0.02, tries=8):
    """Non-overlapping disks in a strip periodic in x (period `width`) that
    starts at y = 0: each disk, in random order, is dropped at `tries`
    uniform x positions and kept at the one where it comes to rest lowest
    (resting where it first touches the floor or a disk below it: random
    ballistic deposition, one pass).  Returns (positions [n, 2], the strip's
    height: the top of the highest disk plus half the gap)."""
    r = np.asarray(radii, np.float64)
    n = len(r)
    W = float(width)
    if 2 * r.max() + gap > W:
        raise ValueError("strip too narrow for the largest disk")
    order = rng.permutation(n)
    xs = rng.uniform(0.0, W, (n, tries))
    px = np.empty(n)
    py = np.empty(n)
    pr = np.empty(n)
    for m, i in enumerate(order):
        ri = r[i]
        x, y = xs[m, 0], ri + 0.5 * gap
        if m:
            dx = px[:m][None, :] - xs[m][:, None]
            dx -= W * np.round(dx / W)
            reach = ri + pr[:m] + gap
            near = np.abs(dx) < reach[None, :]
            top = np.where(near, py[:m][None, :] +
                           np.sqrt(np.maximum(reach[None, :] ** 2 - dx ** 2, 0.0)), 0.0)
            rest = np.maximum(top.max(1), ri + 0.5 * gap)
            k = int(np.argmin(rest))
            x, y = xs[m, k], float(rest[k])
        px[m], py[m], pr[m] = x, y, ri
    pos = np.empty((n, 2))
    pos[order, 0] = px
    pos[order, 1] = py
    return pos, float(np.max(py + pr) + 0.5 * gap)


@dataclass
class Crossing:
    E: float                 # x and z period (um)
    H: float                 # y period (um)
    pq: tuple                # direction (p, 0, q) of bundle B
    angle: float             # crossing angle (degrees)
    a_xy: np.ndarray         # [na, 2] bundle A: x in [0, E), y in [0, H/2)
    a_r: np.ndarray
    b_uy: np.ndarray         # [nb, 2] bundle B: u in [0, P), y in [0, H/2)
    b_r: np.ndarray
    ngon: int
    intra_fraction: float    # polygon area / voxel area, per slab

    @property
    def n_axons(self):
        return len(self.a_r) + len(self.b_r)

    @property
    def period_b(self):
        p, q = self.pq
        return self.E / np.hypot(p, q)

    @property
    def space_scale(self):
        return max(self.E, self.H) / 2.0


def make_crossing(n_axons, mean_radius, angle_deg, rng, *, fraction=None, gamma_k=3.0,
                  ngon=32, gap=0.02, min_period=24.0):
    """Draws radii (Gamma(k) with the given mean, capped at 2.5x the mean),
    sizes the voxel and packs both bundles.  fraction=None deposits the disks
    (deposit_strip, one pass, about 0.5 of the area); a fraction packs each
    slab to it by push-apart (pack_strip, slower)."""
    (p, q), angle = lattice_direction(angle_deg)
    n = np.hypot(p, q)
    r = np.minimum(S.gamma_radii(int(n_axons), gamma_k, mean_radius, rng), 2.5 * mean_radius)
    rmax = float(r.max())
    # bundle B's strip period E/n must hold several of the largest disks
    E = max(min_period, 6.0 * n * rmax)
    P = E / n
    nb = max(1, int(round(n_axons / (n + 1.0))))
    na = max(1, int(n_axons) - nb)
    perm = rng.permutation(len(r))
    ra, rb = r[perm[:na]], r[perm[na:na + nb]]
    poly = 0.5 * ngon * np.sin(2 * np.pi / ngon)
    area_a, area_b = poly * np.sum(ra ** 2), poly * np.sum(rb ** 2)
    if fraction is None:
        pa, ha = deposit_strip(ra, E, rng, gap=gap)
        pb, hb = deposit_strip(rb, P, rng, gap=gap)
        h = max(ha, hb)
    else:
        h = max(area_a / E, area_b / P) / fraction
        h = max(h, 2.0 * rmax + 4 * gap)
        pa, _ = pack_strip(ra, E, h, rng, gap=gap)
        pb, _ = pack_strip(rb, P, h, rng, gap=gap)
    frac = min(area_a / (E * h), area_b / (P * h))
    return Crossing(E=float(E), H=float(2 * h), pq=(p, q), angle=float(angle), a_xy=pa, a_r=ra,
                    b_uy=pb, b_r=rb, ngon=ngon, intra_fraction=float(frac))


def _prism(c0, c1, R, eu, ev, ngon):
    """Vertices [2*ngon, 3] and triangles [2*ngon, 3] (local indices) of an
    open prism from centre c0 to c1 with radius R in the plane (eu, ev)."""
    ang = 2 * np.pi * np.arange(ngon) / ngon
    ring = R * (np.cos(ang)[:, None] * eu[None, :] + np.sin(ang)[:, None] * ev[None, :])
    V = np.concatenate([c0[None, :] + ring, c1[None, :] + ring])
    b = np.arange(ngon)
    t = ngon + b
    bn, tn = np.roll(b, -1), np.roll(t, -1)
    T = np.concatenate([np.stack([b, bn, tn], 1), np.stack([b, tn, t], 1)])
    return V, T


def crossing_mesh(cx: Crossing, halo):
    """Mesh of both bundles with every image that reaches the halo box
    [-E/2 - halo, E/2 + halo] x [-H/2 - halo, H/2 + halo] x (as x).
    Returns (vertices normalised, triangles, tri_info, space_scale)."""
    E, H, ngon = cx.E, cx.H, cx.ngon
    s = cx.space_scale
    hx = E / 2 + halo
    hy = H / 2 + halo
    verts, tris, fibs = [], [], []
    nv = 0

    def add(c0, c1, R, eu, ev, f):
        nonlocal nv
        V, T = _prism(c0, c1, R, eu, ev, ngon)
        verts.append(V)
        tris.append(T + nv)
        fibs.append(np.full(len(T), f, np.uint32))
        nv += len(V)

    ex, ey, ez = np.eye(3)
    # bundle A: along z, slab y in [-H/2, 0)
    for f, ((x, y), R) in enumerate(zip(cx.a_xy, cx.a_r)):
        X0, Y0 = x - E / 2, y - H / 2
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                X, Y = X0 + ix * E, Y0 + iy * H
                if abs(X) - R > hx or abs(Y) - R > hy:
                    continue
                zlen = hx + R
                add(np.array([X, Y, -zlen]), np.array([X, Y, zlen]), R, ex, ey, f)
    # bundle B: along d, slab y in [0, H/2)
    p, q = cx.pq
    n = np.hypot(p, q)
    d = np.array([p, 0.0, q]) / n
    u = np.array([q, 0.0, -p]) / n
    P = E / n
    na = len(cx.a_r)
    for j, ((uu, y), R) in enumerate(zip(cx.b_uy, cx.b_r)):
        U0 = uu - P / 2
        reach = hx * (abs(p) + abs(q)) / n + R  # |u| of a line that meets the halo box
        kmin = int(np.floor((-reach - U0) / P))
        kmax = int(np.ceil((reach - U0) / P))
        for k in range(kmin, kmax + 1):
            Uk = U0 + k * P
            for iy in (-1, 0, 1):
                Y = y + iy * H
                if abs(Y) - R > hy:
                    continue
                c = Uk * u + Y * ey
                # chord of the line through the halo box grown by R (x and z)
                t0, t1 = -np.inf, np.inf
                lim = hx + R
                for a in (0, 2):
                    if abs(d[a]) < 1e-12:
                        if abs(c[a]) > lim:
                            t0, t1 = 1.0, -1.0
                        continue
                    ta, tb = (-lim - c[a]) / d[a], (lim - c[a]) / d[a]
                    t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
                if not t0 < t1:
                    continue
                add(c + t0 * d, c + t1 * d, R, u, ey, na + j)
    V = (np.concatenate(verts) / s).astype(np.float32)
    T = np.concatenate(tris).astype(np.uint32)
    F = np.concatenate(fibs)
    I = np.zeros_like(T)
    I[:, 0] = F
    I[:, 2] = T[:, 0]
    return V, T, I, s


def voxel_bounds(cx: Crossing):
    """The periodic voxel [lo, hi) per axis, normalised."""
    s = cx.space_scale
    hi = np.array([cx.E / 2, cx.H / 2, cx.E / 2]) / s
    return -hi, hi


def grid_params(cx: Crossing, halo, cell_um, eps_rel=4e-3):
    """Exact-membership grid over the halo box with cubic cells of cell_um."""
    s = cx.space_scale
    ext = np.array([cx.E, cx.H, cx.E]) / 2 + halo
    dims = tuple(int(np.ceil(2 * e / cell_um)) for e in ext)
    c = tuple(float(2 * e / s / k) for e, k in zip(ext, dims))
    lo = tuple(float(-e / s) for e in ext)
    return dict(dims=dims, lo=lo, cell=c, eps=float(eps_rel * min(c)))


def inside(cx: Crossing, pts_um):
    """Reference labels of points (micrometres, voxel-centred, inside the
    voxel): fibre index of the polygon holding each point, or -1, and the
    distance to the nearest wall (polygon edge).  For tests."""
    E, H, ngon = cx.E, cx.H, cx.ngon
    p, q = cx.pq
    n = np.hypot(p, q)
    P = E / n
    u = np.array([q, 0.0, -p]) / n
    fib = np.full(len(pts_um), -1)
    dist = np.full(len(pts_um), np.inf)
    ang = 2 * np.pi * np.arange(ngon) / ngon
    half = np.pi / ngon
    na = len(cx.a_r)
    for i, (x, y, z) in enumerate(pts_um):
        if y < 0:  # slab A: transverse (x, y), periodic in x
            cand = [(f, cx.a_xy[f, 0] - E / 2, cx.a_xy[f, 1] - H / 2, cx.a_r[f]) for f in range(na)]
            tx, ty, per = x, y, E
        else:
            tu = x * u[0] + z * u[2]
            cand = [(na + j, cx.b_uy[j, 0] - P / 2, cx.b_uy[j, 1], cx.b_r[j])
                    for j in range(len(cx.b_r))]
            tx, ty, per = tu, y, P
        for f, X, Y, R in cand:
            dx = tx - X
            dx -= per * np.round(dx / per)
            for dy in (ty - Y, ty - Y + H, ty - Y - H):
                rho = np.hypot(dx, dy)
                if rho > R + 1.0:
                    continue
                # distance to the regular polygon's boundary along the
                # sector's apothem
                phi = np.mod(np.arctan2(dy, dx), 2 * np.pi)
                k = np.floor(phi / (2 * np.pi / ngon))
                mid = (k + 0.5) * 2 * np.pi / ngon
                apo = R * np.cos(half)
                proj = rho * np.cos(phi - mid)
                dist[i] = min(dist[i], abs(proj - apo))
                if proj < apo:
                    fib[i] = f
    del ang
    return fib, dist

"""Synthetic axon substrates for the benchmark configurations.

Axons are z-aligned N-gon prisms (the shape dSimGetFiberTriangles.m builds:
2N side triangles per segment, optional fan caps) packed without overlap in a
periodic square voxel.  Positions come from random placement followed by
iterative overlap removal (pairwise push-apart with the periodic minimum
image), which reaches the dense fractions random sequential adsorption cannot.
Every axon that reaches into the grid halo around the voxel is copied by the
voxel period, with the same fibre index, so a spin near a face sees the
neighbouring side's membranes.

All coordinates returned are normalised: micrometres / space_scale, with the
voxel spanning [-1, 1] on every axis (space_scale = half the voxel side).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import cKDTree


@dataclass
class Axons:
    centers: np.ndarray      # [n, 2] micrometres, in [0, L)
    radii: np.ndarray        # [n] micrometres (polygon circumradius)
    voxel: float             # L, micrometres
    intra_fraction: float    # polygon area / voxel area
    info: dict = field(default_factory=dict)


def pack_disks(radii, voxel, rng, gap=0.02, max_iter=20000, tol=1e-9):
    """Non-overlapping periodic packing of disks with the given radii."""
    r = np.asarray(radii, np.float64)
    n = len(r)
    L = float(voxel)
    pos = rng.uniform(0.0, L, (n, 2))
    # place large disks first by sorting; push-apart handles the rest
    rmax = r.max()
    for it in range(max_iter):
        tree = cKDTree(pos, boxsize=L)
        pairs = tree.query_pairs(2 * rmax + gap, output_type="ndarray")
        if len(pairs) == 0:
            return pos, it
        i, j = pairs[:, 0], pairs[:, 1]
        d = pos[j] - pos[i]
        d -= L * np.round(d / L)
        dist = np.hypot(d[:, 0], d[:, 1])
        need = r[i] + r[j] + gap
        ov = need - dist
        m = ov > tol
        if not m.any():
            return pos, it
        i, j, d, dist, ov = i[m], j[m], d[m], dist[m], ov[m]
        dist = np.maximum(dist, 1e-12)
        u = d / dist[:, None]
        # move both disks apart, shared by area so big disks move less
        wi = r[j] ** 2 / (r[i] ** 2 + r[j] ** 2)
        wj = 1.0 - wi
        step = (ov * 0.5 + 1e-9)[:, None] * 1.02
        delta = np.zeros_like(pos)
        np.add.at(delta, i, -u * step * 2 * wi[:, None])
        np.add.at(delta, j, u * step * 2 * wj[:, None])
        pos = np.mod(pos + delta, L)
    raise RuntimeError(f"packing did not converge in {max_iter} iterations")


def gamma_radii(n, k, mean, rng, rmin=0.1):
    r = rng.gamma(k, mean / k, n)
    return np.maximum(r, rmin)


def make_axons(n_axons, k, mean_radius, rng, *, voxel=None, fraction=None, ngon=64, gap=0.02):
    """Draw Gamma radii and pack them.  Give either the voxel side or the
    target intra-axonal area fraction (the voxel is then sized to match)."""
    r = gamma_radii(n_axons, k, mean_radius, rng)
    poly = 0.5 * ngon * np.sin(2 * np.pi / ngon)      # N-gon area / R^2
    area = poly * np.sum(r ** 2)
    if voxel is None:
        voxel = float(np.sqrt(area / fraction))
    pos, iters = pack_disks(r, voxel, rng, gap=gap)
    return Axons(pos, r, float(voxel), float(area / voxel ** 2),
                 dict(pack_iterations=iters, ngon=ngon))


def prism_mesh(axons: Axons, *, ngon, z_range, caps=False, halo=0.0):
    """Mesh of z-aligned N-gon prisms with periodic copies.

    z_range: (z0, z1) micrometres relative to the voxel centre.
    halo: micrometres beyond each voxel face that must see correct geometry.
    Returns (vertices[nv,3] normalised, triangles[nt,3], tri_info[nt,3],
    space_scale)."""
    L = axons.voxel
    s = L / 2.0
    verts, tris, info = [], [], []
    nv = 0
    ang = 2 * np.pi * np.arange(ngon) / ngon
    cs, sn = np.cos(ang), np.sin(ang)
    z0, z1 = z_range
    for f, ((cx, cy), R) in enumerate(zip(axons.centers, axons.radii)):
        # voxel-centred coordinates, then every periodic image touching the halo box
        bx, by = cx - s, cy - s
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                x, y = bx + ix * L, by + iy * L
                if abs(x) - R > s + halo or abs(y) - R > s + halo:
                    continue
                ring = np.stack([x + R * cs, y + R * sn], 1)
                bot = np.column_stack([ring, np.full(ngon, z0)])
                top = np.column_stack([ring, np.full(ngon, z1)])
                verts += [bot, top]
                b = nv + np.arange(ngon)
                t = nv + ngon + np.arange(ngon)
                bn, tn = np.roll(b, -1), np.roll(t, -1)
                tris.append(np.stack([b, bn, tn], 1))
                tris.append(np.stack([b, tn, t], 1))
                nv += 2 * ngon
                if caps:
                    verts.append(np.array([[x, y, z0], [x, y, z1]]))
                    cb, ct = nv, nv + 1
                    tris.append(np.stack([np.full(ngon, cb), bn, b], 1))
                    tris.append(np.stack([np.full(ngon, ct), t, tn], 1))
                    nv += 2
    V = (np.concatenate(verts) / s).astype(np.float32)
    T = np.concatenate(tris).astype(np.uint32)
    # fibre index per triangle, membrane 0 (axon wall), first vertex
    fib = []
    per = 2 * ngon + (2 * ngon if caps else 0)
    k = 0
    for f, ((cx, cy), R) in enumerate(zip(axons.centers, axons.radii)):
        bx, by = cx - s, cy - s
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                x, y = bx + ix * L, by + iy * L
                if abs(x) - R > s + halo or abs(y) - R > s + halo:
                    continue
                fib.append(np.full(per, f))
    fib = np.concatenate(fib).astype(np.uint32)
    assert len(fib) == len(T)
    I = np.zeros_like(T)
    I[:, 0] = fib
    I[:, 2] = T[:, 0]
    return V, T, I, s


@dataclass
class Config:
    name: str
    n_spins: int
    n_axons: int
    gamma_k: float
    mean_radius: float
    ngon: int
    voxel: float | None
    fraction: float | None
    length: float | None        # prism length (None: spans the periodic z period)
    caps: bool
    D: float                    # um^2/ms, all compartments
    dt: float                   # ms
    t2_intra: float
    t2_extra: float
    kappa: float                # permeability um/ms
    delta: float
    Delta: float
    TE: float
    b_shells: tuple
    dirs_per_shell: tuple
    seed: int


# Config interpretations (sizes as stated for each benchmark configuration):
#  small: 16 axons, Gamma(k=4, mean 2 um), 32-gon prisms 100 um long with caps,
#         20 um periodic voxel, D=2, dt=5 us, impermeable, T2 off,
#         PGSE 10/40/60 ms, 6 directions at b=1 ms/um^2 (+ b=0), 1e4 spins.
#  scale: 1e6 spins, 2000 axons Gamma(k=3, mean 1 um), 64-gon open prisms,
#         voxel sized for intra-axonal fraction 0.7 (about 109 um), kappa
#         0.01 um/ms, T2 70/90 ms, 64 directions on b=1/2/3 ms/um^2 shells
#         (21/21/22), PGSE 10/40/60 ms.
SMALL = Config("small-16axon-1e4", 10_000, 16, 4.0, 2.0, 32, 20.0, None, 100.0, True, 2.0, 0.005,
               np.inf, np.inf, 0.0, 10.0, 40.0, 60.0, (1.0,), (6,), 1)
SCALE = Config("scale-2000axon-1e6", 1_000_000, 2000, 3.0, 1.0, 64, None, 0.7, None, False, 2.0,
               0.005, 70.0, 90.0, 0.01, 10.0, 40.0, 60.0, (1.0, 2.0, 3.0), (21, 21, 22), 2)


def kappa_to_probability(kappa, D, dt):
    """Per-hit transmission probability matching a membrane permeability.

    A membrane of permeability kappa carries flux J = kappa * c.  In one step
    of the walk the normal displacement X ~ N(0, sigma^2), sigma^2 = 2 D dt;
    spins within reach cross with probability p, so the walk's flux is
    J = c * p * E[max(X, 0)] / dt = c * p * sigma / (sqrt(2 pi) dt).
    Equating gives p = kappa * sqrt(2 pi) * dt / sigma = kappa * sqrt(pi dt / D)."""
    return float(kappa * np.sqrt(np.pi * dt / D))


def fibonacci_directions(n, offset=0.5):
    """Unit vectors spread over the sphere (golden-angle spiral)."""
    k = np.arange(n) + offset
    z = 1 - 2 * k / n
    r = np.sqrt(np.maximum(0.0, 1 - z * z))
    phi = np.pi * (3 - np.sqrt(5)) * k
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def gradient_amplitude(b, delta, Delta, gamma=42576.0):
    """|G| in mT/m for a PGSE b-value in ms/um^2 (b formula of
    diffusion.cpp:317)."""
    g = 2 * np.pi * gamma * 1e-9 * delta
    return np.sqrt(b / (Delta - delta / 3.0)) / g


def protocol(cfg: Config):
    """Gradient rows [n, 6]: delta Delta readout gx gy gz (sim.grads layout),
    b=0 first."""
    rows = [[cfg.delta, cfg.Delta, cfg.TE - cfg.Delta - cfg.delta, 0.0, 0.0, 0.0]]
    for b, nd in zip(cfg.b_shells, cfg.dirs_per_shell):
        G = gradient_amplitude(b, cfg.delta, cfg.Delta)
        for d in fibonacci_directions(nd):
            rows.append([cfg.delta, cfg.Delta, cfg.TE - cfg.Delta - cfg.delta, *(G * d)])
    return np.array(rows, np.float32)


def pack(cfg: Config) -> Axons:
    """The configuration's axon packing (seeded, deterministic)."""
    rng = np.random.default_rng(cfg.seed)
    return make_axons(cfg.n_axons, cfg.gamma_k, cfg.mean_radius, rng, voxel=cfg.voxel,
                      fraction=cfg.fraction, ngon=cfg.ngon)


def build(cfg: Config, halo_cells=2, cell_um=None, axons: Axons | None = None):
    """Mesh + grid parameters for a configuration (axons: reuse a packing)."""
    ax = axons if axons is not None else pack(cfg)
    L = ax.voxel
    s = L / 2.0
    sigma_um = np.sqrt(2 * cfg.D * cfg.dt)
    if cell_um is None:
        cell_um = max(1.0, 4 * sigma_um)
    halo = max(halo_cells * cell_um, 12 * sigma_um)
    if cfg.length is None:
        zr = (-(s + halo), s + halo)
    else:
        zr = (-cfg.length / 2, cfg.length / 2)
    V, T, I, space_scale = prism_mesh(ax, ngon=cfg.ngon, z_range=zr, caps=cfg.caps, halo=halo)
    lo_xy = -(s + halo) / s
    nxy = int(np.ceil(2 * (s + halo) / cell_um))
    cxy = 2 * (s + halo) / s / nxy
    grid = dict(dims=(nxy, nxy, 1), lo=(lo_xy, lo_xy, lo_xy), cell=(cxy, cxy, -2 * lo_xy),
                eps=1e-5 * cxy)
    sd = np.float32(sigma_um / space_scale)
    p = kappa_to_probability(cfg.kappa, cfg.D, cfg.dt)
    t2 = (cfg.t2_extra, cfg.t2_intra, cfg.t2_intra, cfg.t2_intra)
    xmax = float(np.max(np.abs(V[:, 0]))) if len(V) else 1.0
    return dict(axons=ax, vertices=V, triangles=T, tri_info=I, space_scale=space_scale, grid=grid,
                std_dev=(sd,) * 4, t2=t2, permeability=p, xmax=max(1.0, xmax), halo_um=halo,
                cell_um=cell_um, protocol=protocol(cfg), cfg=cfg)

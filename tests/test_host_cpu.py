"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared entry point, and the product's substrate preparation (triangle
helpers, reference grid, AABB grid) is bit-identical to the oracle's."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D


def test_library_exports_every_declared_symbol():
    names = D.exported_names()
    assert len(names) >= 25
    L = C.CDLL(D.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert D.lib().dsim_abi_version() == 1


def test_no_device_is_reported_cleanly():
    if D.device_count() > 0:
        pytest.skip("a device is visible")
    sub = D.Substrate(np.zeros((3, 3), np.float32), [[0, 1, 2]], [[0, 0, 0]])
    with pytest.raises(D.DsimError) as e:
        D.Simulation(sub, 10, dt=0.005, std_dev=(0.01,) * 4)
    assert e.value.code == -8


def test_bad_mesh_rejected():
    with pytest.raises(D.DsimError) as e:
        D.Substrate(np.zeros((3, 3), np.float32), [[0, 1, 5]], [[0, 0, 0]])
    assert e.value.code == -1
    with pytest.raises(D.DsimError):
        D.Substrate(np.zeros((3, 3), np.float32), [[0, 1, 2]], [[0, 7, 0]])


MESHES = [("sim.fibers_1_r004_f09", 60.0), ("sim.fibers_1_r004_f09", 3.0),
          ("sim.fibers_1_r0064_f05", 4.0), ("sim.triangles_curved", 1.0)]


@pytest.mark.parametrize("name,scale", MESHES)
def test_reference_grid_and_helpers_match_oracle(ref_inputs, name, scale):
    m = O.parse_mesh(os.path.join(ref_inputs, name), scale)
    sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], grid="reference", n_cubes=50)
    np.testing.assert_array_equal(sub.helpers(), O.tri_helpers(m["vertices"], m["trgls"]))
    start, tri, maxc = sub.grid()
    cnt, lst, omax = O.populate_cubes(m["vertices"], m["trgls"], 50)
    assert maxc == omax
    np.testing.assert_array_equal(np.diff(start), cnt)
    for c in np.nonzero(cnt)[0]:
        np.testing.assert_array_equal(tri[start[c]:start[c + 1]], lst[c, :cnt[c]])


@pytest.mark.parametrize("name,scale", MESHES[1:])
def test_aabb_grid_matches_oracle(ref_inputs, name, scale):
    m = O.parse_mesh(os.path.join(ref_inputs, name), scale)
    dims, lo, cell = (23, 19, 11), (-1.2, -1.1, -1.3), (2.4 / 23, 2.2 / 19, 2.6 / 11)
    sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=dims, lo=lo,
                      cell=cell, eps=1e-5)
    start, tri, maxc = sub.grid()
    cnt, lst, omax = O.populate_aabb(m["vertices"], m["trgls"], dims, lo, cell, 1e-5)
    assert maxc == omax
    np.testing.assert_array_equal(np.diff(start), cnt)
    for c in np.nonzero(cnt)[0]:
        np.testing.assert_array_equal(tri[start[c]:start[c + 1]], lst[c, :cnt[c]])


def _point_triangle_distance(p, a, b, c):
    """Exact Euclidean distance from points p[n,3] to triangle abc (float64)."""
    p = p.astype(np.float64)
    a, b, c = (np.asarray(v, np.float64) for v in (a, b, c))
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    nn = n @ n
    # projection onto the plane and its barycentrics
    t = ((p - a) @ n) / nn
    q = p - t[:, None] * n
    v0, v1, v2 = ab, ac, q - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    inside = (v >= 0) & (w >= 0) & (v + w <= 1)
    best = np.where(inside, np.abs(t) * np.sqrt(nn), np.inf)
    for s0, s1 in ((a, b), (b, c), (c, a)):
        e = s1 - s0
        u = np.clip(((p - s0) @ e) / (e @ e), 0, 1)
        best = np.minimum(best, np.linalg.norm(p - (s0 + u[:, None] * e), axis=1))
    return best


@pytest.mark.parametrize("name,scale", [("sim.fibers_1_r004_f09", 3.0), ("sim.triangles_curved", 1.0)])
def test_clearance_is_a_distance_lower_bound(ref_inputs, name, scale):
    m = O.parse_mesh(os.path.join(ref_inputs, name), scale)
    dims, lo, cell = (31, 27, 9), (-1.2, -1.2, -1.2), (2.4 / 31, 2.4 / 27, 2.4 / 9)
    sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=dims, lo=lo,
                      cell=cell, eps=1e-5)
    clr = sub.clearance()
    assert clr.shape == (int(np.prod(dims)),) and (clr >= 0).all()
    rng = np.random.default_rng(1)
    V, T = m["vertices"], m["trgls"]
    cells = rng.choice(len(clr), 400, replace=False)
    lo_, ce = np.array(lo, np.float64), np.array(cell, np.float64)
    tight = 0
    for c in cells:
        ix, iy, iz = c % dims[0], (c // dims[0]) % dims[1], c // (dims[0] * dims[1])
        pts = lo_ + (np.array([ix, iy, iz]) + rng.uniform(0, 1, (64, 3))) * ce
        d = np.full(len(pts), np.inf)
        for t in T:
            d = np.minimum(d, _point_triangle_distance(pts, *V[t]))
        assert (clr[c] <= d + 1e-12).all(), (c, clr[c], d.min())
        tight += clr[c] > 0.5 * d.min()
    assert tight > 40  # the bound is useful, not just zero


def test_autotuner_choice_and_launch_length():
    # near-ties (within 1 %) go to the longest reorder interval, then the
    # largest cell; a clearly faster candidate wins outright
    from dsim_amd import autotune as AT
    rs = [dict(cell_um=0.15, sort_interval=512, rate=1.650e10),
          dict(cell_um=0.2, sort_interval=1024, rate=1.640e10),
          dict(cell_um=0.15, sort_interval=1024, rate=1.645e10),
          dict(cell_um=0.25, sort_interval=1024, rate=1.500e10)]
    b = AT.pick(rs)
    assert (b["cell_um"], b["sort_interval"]) == (0.2, 1024)
    rs[0]["rate"] = 1.80e10
    b = AT.pick(rs)
    assert (b["cell_um"], b["sort_interval"]) == (0.15, 512)
    assert [AT.launch_steps(s) for s in (0, 128, 256, 512, 1024)] == [1024, 256, 256, 512, 1024]

"""Crossing-bundle substrates on the GPU (the batched configuration's
geometry: two bundles of open prisms at lattice angles in a non-cubic
periodic voxel, exact-membership grids): device grid build against the host
build, labels against the oracle along the substrate's ray axis, injected
walks bit-exact against the oracle, and many crossing substrates advanced by
the same batched launches (served bits equal the plain path's, signals agree
with the oracle's walk of the same spins)."""
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import crossing as X
from dsim_amd import lib as D
from dsim_amd import serve as SV

pytestmark = pytest.mark.gpu

GAMMA = 42576.0
DT = 0.005
HALO = 1.5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _case(angle, n=60, r=1.0, seed=3, cell_um=0.75):
    cx = X.make_crossing(n, r, angle, np.random.default_rng(seed))
    V, T, I, s = X.crossing_mesh(cx, HALO)
    g = X.grid_params(cx, HALO, cell_um)
    return cx, V, T, I, s, g


def _oracle_world(cx, V, T, I, s, **kw):
    # the oracle's own grid (box overlap and the triangle's plane through the
    # cell, 1 um cells): the nearest-hit rule does not depend on it
    lo, hi = X.voxel_bounds(cx)
    ext = np.array([cx.E, cx.H, cx.E]) / 2 / s + HALO / s
    dims = tuple(int(np.ceil(2 * e / (1.0 / s))) for e in ext)
    return O.World(V, T, I, grid="aabb_plane", dims=dims, lo=tuple(-ext),
                   cell=tuple(2 * ext / np.array(dims)), eps=1e-6, boundary=1, test=1,
                   traversal=1, transitions=1, voxel_lo=tuple(lo), voxel_hi=tuple(hi), **kw)


@pytest.mark.parametrize("angle", [45, 90])
def test_tribox_device_build_equals_host_build(angle):
    _say(f"tribox device build, angle {angle}")
    cx, V, T, I, s, g = _case(angle)
    h = D.Substrate(V, T, I, grid="tribox", **g)
    d = D.Substrate(V, T, I, grid="tribox", device=True, **g)
    sh, th, mh = h.grid()
    sd, td, md = d.grid()
    assert np.array_equal(sh, sd) and np.array_equal(th, td) and mh == md
    np.testing.assert_array_equal(h.clearance(), d.clearance())
    np.testing.assert_array_equal(h.helpers(), d.helpers())
    assert h.label_axis() == d.label_axis() == 1


@pytest.mark.parametrize("angle", [45, 90])
@pytest.mark.parametrize("grid", ["tribox", "aabb"])
def test_crossing_labels_match_oracle(angle, grid):
    _say(f"crossing labels, angle {angle}, {grid} grid")
    cx, V, T, I, s, g = _case(angle)
    sub = D.Substrate(V, T, I, grid=grid, device=True, **g)
    axis = sub.label_axis()
    n = 6000
    lo, hi = X.voxel_bounds(cx)
    sim = D.Simulation(sub, n, dt=DT, std_dev=(0.01,) * 4, mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, voxel=(lo, hi))
    pos = np.random.default_rng(5).uniform(lo, hi, (n, 3)).astype(np.float32)
    w = _oracle_world(cx, V, T, I, s)
    mx = np.abs(V).max(0)
    comp, fib, tries, amb = w.find_spins(pos, *mx, ray_axis=axis)
    sim.set_state(pos=pos)
    gamb = sim.label(*mx)
    st = sim.get_state()
    assert gamb == amb
    np.testing.assert_array_equal(st["comp"], comp)
    np.testing.assert_array_equal(st["fiber"], fib)
    want, dist = X.inside(cx, pos.astype(np.float64) * s)
    ok = dist > 1e-3 * s
    np.testing.assert_array_equal(np.where(comp == 1, fib.astype(int), -1)[ok], want[ok])


def test_tribox_row_labels_on_z_prisms():
    # +x rays over an exact-membership grid of z-aligned prisms (with images
    # reaching past the grid): the row walk plus the outside list equal the
    # oracle's brute force
    from dsim_amd import substrate as S
    _say("tribox row labels, z prisms")
    ax = S.make_axons(40, 3.0, 1.5, np.random.default_rng(4), fraction=0.55, ngon=32)
    V, T, I, s = S.prism_mesh(ax, ngon=32, z_range=(-0.6 * ax.voxel, 0.6 * ax.voxel), halo=1.0)
    g = dict(dims=(30, 30, 12), lo=(-1.05, -1.05, -1.05), cell=(2.1 / 30, 2.1 / 30, 2.1 / 12),
             eps=1e-5)
    sub = D.Substrate(V, T, I, grid="tribox", device=True, **g)
    assert sub.label_axis() == 0
    n = 8000
    sim = D.Simulation(sub, n, dt=DT, std_dev=(0.01,) * 4, mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC)
    pos = np.random.default_rng(6).uniform(-1, 1, (n, 3)).astype(np.float32)
    w = O.World(V, T, I, test=1, traversal=1, transitions=1, grid="aabb", dims=(1, 1, 1),
                lo=(-2, -2, -2), cell=(4, 4, 4), eps=0.0)
    mx = np.abs(V).max(0)
    comp, fib, tries, amb = w.find_spins(pos, *mx, ray_axis=0)
    sim.set_state(pos=pos)
    assert sim.label(*mx) == amb
    st = sim.get_state()
    np.testing.assert_array_equal(st["comp"], comp)
    np.testing.assert_array_equal(st["fiber"], fib)


@pytest.mark.parametrize("angle", [26.6, 90])
def test_crossing_injected_walk_bit_exact(angle):
    _say(f"crossing injected walk, angle {angle}")
    cx, V, T, I, s, g = _case(angle, cell_um=0.5)
    sd = np.float32(0.25 / s)  # 0.25 um steps: many wall hits
    perm = 0.3
    sub = D.Substrate(V, T, I, grid="tribox", device=True, **g)
    lo, hi = X.voxel_bounds(cx)
    n, iters = 2048, 96
    sim = D.Simulation(sub, n, dt=DT, std_dev=(sd,) * 4, t2=(90.0, 70.0, 70.0, 70.0),
                       permeability=perm, mode=D.MODE_PHYSICAL, boundary=D.BOUNDARY_PERIODIC,
                       voxel=(lo, hi))
    w = _oracle_world(cx, V, T, I, s, std_dev=(sd,) * 4, t2=(90.0, 70.0, 70.0, 70.0), dt=DT,
                      permeability=perm)
    rng = np.random.default_rng(8)
    pos = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    mx = np.abs(V).max(0)
    comp, fib, _, _ = w.find_spins(pos, *mx, ray_axis=sub.label_axis())
    q = np.empty((iters, n, 4), np.float32)
    q[..., :3] = rng.standard_normal((iters, n, 3))
    q[..., 3] = rng.uniform(1e-6, 1.0, (iters, n))
    grad = np.array([0.02, -0.01, 0.03], np.float32)
    pc = np.float32(2 * np.pi * GAMMA * DT)
    st = O.State(pos, comp=comp, fiber=fib)
    tr_o = O.integrate(w, st, iters, rng=O.RNG_INJECT, inject=q, grad=grad, phase_constant=pc,
                       trace=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    tr_g = sim.integrate_injected(grad, pc, q, trace=True)
    gs = sim.get_state()
    hits = int((tr_o[..., 2] > 0).sum())
    _say(f"  {hits} colliding steps, {int((tr_o[..., 2] > 1).sum())} with several hits")
    assert hits > n * iters * 0.02
    # hit triangles, hit counts and labels bit for bit (the cell column
    # counts the grid of each side, which differ)
    assert np.array_equal(tr_g[..., [0, 1, 2, 4]], tr_o[..., [0, 1, 2, 4]])
    assert np.array_equal(gs["pos"], st.pos)
    assert np.array_equal(gs["phase"], st.phase)
    assert np.array_equal(gs["mag"], st.mag)
    assert np.array_equal(gs["comp"], st.comp) and np.array_equal(gs["fiber"], st.fiber)


def _plain(sub, q):
    sim = D.Simulation(sub, q["n"], dt=DT, std_dev=q["std_dev"], t2=q["t2"],
                       permeability=q["permeability"], mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=q["seed"],
                       voxel=q["voxel"])
    sim.init_box(*q["voxel"])
    sim.label(*q["extent"])
    st0 = sim.get_state()
    sim.moment_reset()
    for w, k in q["lobes"]:
        if k:
            sim.integrate_moment(w, k)
    out = sim.signal_moment(q["grads"], q["pc"])
    sim.close()
    return out, st0


def test_many_crossing_substrates_in_one_batched_launch():
    # 16 crossing substrates (40-90 axons, mean radius 0.6-1.5 um, every
    # lattice angle, exact-membership grids built on the device), one request
    # on each: all admitted together and advanced by the same batched
    # launches; every result equals the plain path's bits, and the oracle's
    # walk of the same spins (same Philox streams) gives signals within 3 MC
    # sigma of them
    rng = np.random.default_rng(17)
    subs, reqs, worlds = [], [], []
    for k in range(16):
        angle = float(rng.uniform(0, 90))
        cx, V, T, I, s, g = _case(angle, n=int(rng.integers(40, 90)),
                                  r=float(rng.uniform(0.6, 1.5)), seed=200 + k,
                                  cell_um=float(rng.uniform(0.5, 0.9)))
        sub = D.Substrate(V, T, I, grid="tribox", device=True, **g)
        subs.append(sub)
        sd = np.float32(np.sqrt(2 * 2.0 * DT) / s)
        lo, hi = X.voxel_bounds(cx)
        nl = int(rng.integers(6, 12))
        G = 80.0 * 1e-9 * s
        gr = rng.normal(size=(nl, 3))
        gr = (gr / np.linalg.norm(gr, axis=1, keepdims=True) * G).astype(np.float32)
        reqs.append(dict(n=int(rng.integers(400, 900)), seed=3000 + k, std_dev=(sd,) * 4,
                         t2=(90.0, 70.0, 70.0, 70.0), permeability=0.001,
                         lobes=SV.pgse_lobes(160, 400, 0), grads=gr, voxel=(lo, hi),
                         extent=tuple(np.abs(V).max(0)), pc=np.float32(2 * np.pi * GAMMA * DT)))
        worlds.append(_oracle_world(cx, V, T, I, s, std_dev=(sd,) * 4,
                                    t2=(90.0, 70.0, 70.0, 70.0), dt=DT, permeability=0.001))
    srv = SV.Server(8 << 30, round_steps=100, max_active_spins=1 << 20)
    ids = [srv.add_substrate(sb) for sb in subs]
    tickets = [srv.submit(ids[k], q["n"], seed=q["seed"], dt=DT, std_dev=q["std_dev"],
                          t2=q["t2"], permeability=q["permeability"], lobes=q["lobes"],
                          grads=q["grads"], phase_constant=q["pc"],
                          boundary=D.BOUNDARY_PERIODIC, label_extent=q["extent"],
                          voxel=q["voxel"]) for k, q in enumerate(reqs)]
    srv.run_until_idle()
    stats = []
    for k, (t, q) in enumerate(zip(tickets, reqs)):
        _say(f"  request {k}: {q['n']} spins, {len(q['grads'])} lines")
        got, stt = srv.result(t)
        want, st0 = _plain(subs[k], q)
        np.testing.assert_array_equal(got, want)
        stats.append(stt)
        # the oracle's walk from the same start, same streams
        key = (q["seed"] & 0xFFFFFFFF, q["seed"] >> 32)
        ost = O.State(st0["pos"], comp=st0["comp"], fiber=st0["fiber"], moment=True, phase=None)
        x0 = ost.pos.copy()
        step = 0
        wsum = 0.0
        for wgt, nsteps in q["lobes"]:
            if nsteps:
                O.integrate(worlds[k], ost, nsteps, rng=O.RNG_PHILOX, key=key, step0=step,
                            weight=wgt)
                step += nsteps
                wsum += wgt * nsteps
        osig = O.moment_signal(ost.mag, ost.moment, ost.comp, q["grads"], q["pc"], x0=x0,
                               wsum=wsum)
        n = q["n"]
        sigma = np.sqrt(np.maximum(1.0 - osig[:, 0] ** 2, 1e-4) / (2 * n))
        assert np.all(np.abs(got[:, 0] - osig[:, 0]) < 3 * sigma + 1e-6), (k, got[:, 0], osig[:, 0])
    assert max(s["admit_s"] for s in stats) <= min(s["done_s"] for s in stats)
    srv.close()
    for sb in subs:
        sb.close()

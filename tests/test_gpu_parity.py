"""GPU parity: the HIP path through the C ABI against the oracle on identical
inputs.  Injected step sequences must match bit for bit (hit triangles, cells,
labels, positions, phases, magnitudes); free-running streams must agree within
the stated tolerances."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D

pytestmark = pytest.mark.gpu

GAMMA = 42576.0
DT = 0.005
PC = np.float32(2 * np.pi * GAMMA * DT)  # phaseConstant, spinSystem.cpp:1938


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _mesh(ref_inputs, name, scale):
    return O.parse_mesh(os.path.join(ref_inputs, name), scale)


def _pair(m, *, mode=0, grid="reference", dims=None, lo=None, cell=None, eps=1e-5,
          std=(0.03, 0.03, 0.01, 0.03), t2=(80.0, 70.0, 7.5, 70.0), perm=0.2, boundary=0,
          rng=D.RNG_PHILOX, n=0, seed=1, sort_interval=0, steps_per_launch=0):
    kw = dict(std_dev=std, t2=t2, dt=DT, permeability=perm)
    if grid == "reference":
        w = O.World(m["vertices"], m["trgls"], m["tri_info"], **kw)
        sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"])
    else:
        w = O.World(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=dims, lo=lo,
                    cell=cell, eps=eps, **kw)
        sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=dims,
                          lo=lo, cell=cell, eps=eps)
    if mode == 1:
        w.set(test=1, traversal=1, transitions=1, boundary=boundary)
    sim = D.Simulation(sub, n, dt=DT, std_dev=std, t2=t2, permeability=perm, mode=mode,
                       boundary=boundary, rng=rng, seed=seed, sort_interval=sort_interval,
                       steps_per_launch=steps_per_launch)
    return w, sub, sim


def _spins(w, m, n, seed, box=0.9, plus_x=False):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-box, box, (n, 3)).astype(np.float32)
    comp, fib, _, _ = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"], plus_x=plus_x)
    return pos, comp, fib


def _inject(n, iters, seed):
    rng = np.random.default_rng(seed)
    q = np.empty((iters, n, 4), np.float32)
    q[..., :3] = rng.standard_normal((iters, n, 3)).astype(np.float32)
    q[..., 3] = (rng.integers(0, 2 ** 32, (iters, n), dtype=np.uint64).astype(np.float32) + 1.0) / 4294967296.0
    return q


CASES = [("sim.fibers_1_r004_f09", 6.0), ("sim.fibers_1_r0064_f05", 8.0), ("sim.triangles_curved", 1.0)]


@pytest.mark.parametrize("name,scale", CASES)
def test_injected_reference_mode_bit_exact(ref_inputs, name, scale):
    _say(f"case reference-injected {name} scale={scale}")
    m = _mesh(ref_inputs, name, scale)
    n, iters = 3000, 40
    w, sub, sim = _pair(m, n=n)
    pos, comp, fib = _spins(w, m, n, 7)
    grad = np.array([0.05, -0.02, 0.01], np.float32)
    q = _inject(n, iters, 8)
    st = O.State(pos, comp=comp, fiber=fib)
    tr_o = O.integrate(w, st, iters, rng=O.RNG_INJECT, inject=q, grad=grad, phase_constant=PC,
                       trace=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    tr_g = sim.integrate_injected(grad, PC, q, trace=True)
    g = sim.get_state()
    assert (tr_o[..., 2] > 0).sum() > 500, "too few collisions to be a meaningful check"
    np.testing.assert_array_equal(tr_g, tr_o)
    np.testing.assert_array_equal(g["pos"], st.pos)
    np.testing.assert_array_equal(g["phase"], st.phase)
    np.testing.assert_array_equal(g["mag"], st.mag)
    np.testing.assert_array_equal(g["comp"], st.comp)
    np.testing.assert_array_equal(g["fiber"], st.fiber)


@pytest.mark.parametrize("sort", [0, 16])
@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("name,scale", CASES[:2] + [("sim.triangles_curved", 1.0)])
def test_injected_physical_mode_bit_exact(ref_inputs, name, scale, boundary, sort):
    _say(f"case physical-injected {name} scale={scale} boundary={boundary} sort={sort}")
    m = _mesh(ref_inputs, name, scale)
    n, iters = 3000, 48
    dims, lo, cell = (37, 29, 13), (-1.2, -1.2, -1.2), (2.4 / 37, 2.4 / 29, 2.4 / 13)
    w, sub, sim = _pair(m, mode=1, grid="aabb", dims=dims, lo=lo, cell=cell, boundary=boundary, n=n,
                        sort_interval=sort)
    pos, comp, fib = _spins(w, m, n, 9, plus_x=True)
    q = _inject(n, iters, 10)
    st = O.State(pos, comp=comp, fiber=fib, moment=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    sim.moment_reset()
    tr_o, tr_g = [], []
    for k, wgt in enumerate((1.0, 0.0, -1.0)):
        sl = q[k * 16:(k + 1) * 16]
        tr_o.append(O.integrate(w, st, 16, rng=O.RNG_INJECT, inject=sl, weight=wgt, trace=True))
        tr_g.append(sim.integrate_moment_injected(wgt, sl, trace=True))
    g = sim.get_state()
    mom, x0 = sim.get_moment()
    tr_o, tr_g = np.concatenate(tr_o), np.concatenate(tr_g)
    assert (tr_o[..., 2] > 0).sum() > 500
    np.testing.assert_array_equal(tr_g, tr_o)
    np.testing.assert_array_equal(g["pos"], st.pos)
    np.testing.assert_array_equal(g["mag"], st.mag)
    np.testing.assert_array_equal(g["comp"], st.comp)
    np.testing.assert_array_equal(g["fiber"], st.fiber)
    np.testing.assert_array_equal(mom, st.moment)
    np.testing.assert_array_equal(x0, pos)
    if boundary == 1:
        assert (np.abs(st.pos) <= 1.0).all()


def _grazing_sheet(seed, n_side=6, half=0.6):
    """A flat sheet of 2 n_side^2 triangles in the plane through `origin`
    spanned by R[:, 0], R[:, 1] (R a seeded rotation)."""
    rng = np.random.default_rng(seed)
    a, b, c, d = (lambda v: v / np.linalg.norm(v))(rng.normal(size=4))
    R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                  [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                  [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
    g = np.linspace(-half, half, n_side + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    origin = np.array([0.031, -0.017, 0.0013])
    verts = (np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1) @ R.T + origin).astype(np.float32)
    tris = []
    for i in range(n_side):
        for j in range(n_side):
            v = i * (n_side + 1) + j
            tris += [(v, v + n_side + 1, v + 1), (v + 1, v + n_side + 1, v + n_side + 2)]
    return verts, np.array(tris, np.uint32), R, origin


def test_near_grazing_hits_bit_exact():
    # every spin's first step crosses a seeded sheet within 3e-5..6e-4 rad of its
    # plane, so its hit is decided by the FP64 re-test; the steps after it are
    # ordinary
    _say("case near-grazing sheet")
    verts, tris, R, origin = _grazing_sheet(31)
    m = {"vertices": verts, "trgls": tris, "tri_info": np.zeros((len(tris), 3), np.uint32)}
    n, iters = 4000, 8
    std = np.array([0.03, 0.03, 0.01, 0.03], np.float32)
    w, sub, sim = _pair(m, mode=1, grid="aabb", dims=(16, 16, 16), lo=(-1.2,) * 3,
                        cell=(2.4 / 16,) * 3, n=n, std=tuple(std))
    rng = np.random.default_rng(32)
    xy = rng.uniform(-0.45, 0.45, (n, 2))
    phi = rng.uniform(0, 2 * np.pi, n)
    sin_t = 10.0 ** rng.uniform(-4.5, -3.2, n)
    side = rng.choice([-1.0, 1.0], n)
    lq, f = rng.uniform(0.6, 1.6, n), rng.uniform(0.1, 0.9, n)
    pos0 = (np.c_[xy, np.zeros(n)] @ R.T + origin).astype(np.float32)
    comp0, _, _, _ = w.find_spins(pos0)
    h = -side * f * std[comp0] * lq * sin_t
    pos = (np.c_[xy, h] @ R.T + origin).astype(np.float32)
    comp, fib, _, _ = w.find_spins(pos)
    cos_t = np.sqrt(1 - sin_t ** 2)
    d_loc = np.c_[cos_t * np.cos(phi), cos_t * np.sin(phi), side * sin_t] * lq[:, None]
    q = _inject(n, iters, 33)
    q[0, :, :3] = (d_loc @ R.T).astype(np.float32)
    st = O.State(pos, comp=comp, fiber=fib, moment=True)
    tr_o = O.integrate(w, st, iters, rng=O.RNG_INJECT, inject=q, weight=1.0, trace=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    sim.moment_reset()
    tr_g = sim.integrate_moment_injected(1.0, q, trace=True)
    g = sim.get_state()
    mom, _ = sim.get_moment()
    assert (tr_o[0, :, 2] > 0).sum() > n // 2, "the first steps should mostly hit the sheet"
    np.testing.assert_array_equal(tr_g, tr_o)
    np.testing.assert_array_equal(g["pos"], st.pos)
    np.testing.assert_array_equal(g["mag"], st.mag)
    np.testing.assert_array_equal(g["comp"], st.comp)
    np.testing.assert_array_equal(mom, st.moment)


@pytest.mark.parametrize("name,scale", [("sim.fibers_1_r004_f09", 6.0), ("sim.triangles_curved", 1.0)])
def test_injected_physical_mode_fine_grid_bit_exact(ref_inputs, name, scale):
    # cells of 0.016 against steps of ~0.1 per axis: about a fifth of the cell
    # boxes hold more than the wave's 256-cell list and are searched in
    # several slices
    _say(f"case physical-injected fine grid {name}")
    m = _mesh(ref_inputs, name, scale)
    n, iters = 2000, 24
    dims = (150, 150, 150)
    lo, cell = (-1.2, -1.2, -1.2), (2.4 / 150,) * 3
    w, sub, sim = _pair(m, mode=1, grid="aabb", dims=dims, lo=lo, cell=cell, boundary=1, n=n,
                        std=(0.1,) * 4)
    pos, comp, fib = _spins(w, m, n, 19, plus_x=True)
    q = _inject(n, iters, 20)
    st = O.State(pos, comp=comp, fiber=fib, moment=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    sim.moment_reset()
    tr_o = O.integrate(w, st, iters, rng=O.RNG_INJECT, inject=q, weight=1.0, trace=True)
    tr_g = sim.integrate_moment_injected(1.0, q, trace=True)
    g = sim.get_state()
    mom, _ = sim.get_moment()
    # the box of a step: cells spanned per axis, from the step's extent
    ext = np.floor(np.abs(q[..., :3]) * 0.1 / np.array(cell, np.float32)) + 1
    assert (np.prod(ext, axis=-1) > 256).mean() > 0.1, "too few long boxes to exercise slicing"
    assert (tr_o[..., 2] > 0).sum() > 200
    np.testing.assert_array_equal(tr_g, tr_o)
    np.testing.assert_array_equal(g["pos"], st.pos)
    np.testing.assert_array_equal(g["mag"], st.mag)
    np.testing.assert_array_equal(g["comp"], st.comp)
    np.testing.assert_array_equal(g["fiber"], st.fiber)
    np.testing.assert_array_equal(mom, st.moment)


@pytest.mark.parametrize("name,scale", CASES)
def test_labels_match_oracle(ref_inputs, name, scale):
    _say(f"case labels {name} scale={scale}")
    m = _mesh(ref_inputs, name, scale)
    n = 5000
    for mode in (0, 1):
        w, sub, sim = _pair(m, n=n, mode=mode)
        rng = np.random.default_rng(3 + mode)
        pos = rng.uniform(-0.95, 0.95, (n, 3)).astype(np.float32)
        comp, fib, tries, amb = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"], plus_x=mode == 1)
        sim.set_state(pos=pos)
        gamb = sim.label(m["xmax"], m["ymax"], m["zmax"])
        g = sim.get_state()
        assert gamb == amb
        np.testing.assert_array_equal(g["comp"], comp)
        np.testing.assert_array_equal(g["fiber"], fib)
        assert len(np.unique(comp)) >= 2


ROW_GRIDS = [
    # fine cells over the whole mesh, no margin
    dict(dims=(60, 50, 40), lo=(-1.2,) * 3, cell=(0.04, 0.048, 0.06), eps=0.0),
    # part of the mesh: triangles leaving the grid box take the outside list,
    # spins outside it the tiled pass over all triangles
    dict(dims=(17, 11, 7), lo=(-0.3, -0.2, -0.5), cell=(0.03, 0.04, 0.05), eps=2e-5),
    # a box reaching past 1.1 xmax: spins there cast their ray towards -x
    dict(dims=(20, 20, 20), lo=(-2.0,) * 3, cell=(0.2,) * 3, eps=1e-5),
]


@pytest.mark.parametrize("name,scale", CASES)
def test_grid_row_labels_match_oracle(ref_inputs, name, scale):
    # +x rays over AABB grids go through the grid-row labeller first; the
    # oracle tests every triangle
    _say(f"case grid-row labels {name} scale={scale}")
    m = _mesh(ref_inputs, name, scale)
    n = 8000
    for k, g in enumerate(ROW_GRIDS):
        w, sub, sim = _pair(m, n=n, mode=1, grid="aabb", **g)
        rng = np.random.default_rng(40 + k)
        pos = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
        pos[: n // 2] *= np.float32(0.6)   # half of them near the fibres
        comp, fib, tries, amb = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"], plus_x=True)
        sim.set_state(pos=pos)
        gamb = sim.label(m["xmax"], m["ymax"], m["zmax"])
        st = sim.get_state()
        sim.close()
        assert gamb == amb, (k, gamb, amb)
        np.testing.assert_array_equal(st["comp"], comp)
        np.testing.assert_array_equal(st["fiber"], fib)
        assert len(np.unique(comp)) >= 2


@pytest.mark.parametrize("n", [300, 3000])
def test_grid_row_label_recasts_match_oracle(ref_inputs, n):
    # spins level with mesh vertices: their +x rays run through triangle edges,
    # are re-cast with the jitter, and with a low try limit end ambiguous; a
    # queue of 300 is re-cast one spin per block, one of 3000 by the tiled
    # kernel over the queue
    _say(f"case grid-row label re-casts n={n}")
    m = _mesh(ref_inputs, "sim.fibers_1_r004_f09", 6.0)
    V = np.asarray(m["vertices"], np.float32)
    rng = np.random.default_rng(77 + n)
    k = rng.integers(0, len(V), n)
    # y of a vertex: the ray meets the prism's vertical edge through it
    pos = np.c_[rng.uniform(-1.0, 0.0, n), V[k, 1], rng.uniform(-0.9, 0.9, n)].astype(np.float32)
    for max_tries in (0, 2, 20):
        w, sub, sim = _pair(m, n=n, mode=1, grid="aabb", dims=(60, 50, 40), lo=(-1.2,) * 3,
                            cell=(0.04, 0.048, 0.06), eps=1e-5)
        comp, fib, tries, amb = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"],
                                             max_tries=max_tries, plus_x=True)
        sim.set_state(pos=pos)
        gamb = sim.label(m["xmax"], m["ymax"], m["zmax"], max_tries=max_tries)
        st = sim.get_state()
        sim.close()
        assert gamb == amb, (max_tries, gamb, amb)
        np.testing.assert_array_equal(st["comp"], comp)
        np.testing.assert_array_equal(st["fiber"], fib)
        if max_tries == 0:
            assert amb > n // 10, amb      # most rays touch an edge
        assert (tries > 0).sum() > n // 10


def test_signal_matches_oracle(ref_inputs):
    _say("case signal")
    m = _mesh(ref_inputs, "sim.fibers_1_r004_f09", 6.0)
    n = 20000
    w, sub, sim = _pair(m, n=n)
    rng = np.random.default_rng(12)
    pos = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    comp = rng.integers(0, 4, n).astype(np.uint8)
    mag = rng.uniform(0.2, 1.0, n).astype(np.float32)
    phase = rng.uniform(-40, 40, n).astype(np.float32)
    sim.set_state(pos=pos, comp=comp, mag=mag, phase=phase)
    np.testing.assert_allclose(sim.signal(), O.mr_signal(mag, phase, comp), rtol=1e-11, atol=1e-14)


def test_moment_signal_matches_oracle():
    _say("case moment-signal")
    V = np.zeros((3, 3), np.float32) + 5
    n = 30000
    w, sub, sim = _pair({"vertices": V, "trgls": [[0, 1, 2]], "tri_info": [[0, 0, 0]]}, n=n,
                        mode=1, grid="aabb", dims=(4, 4, 4), lo=(-1, -1, -1), cell=(0.5,) * 3)
    sim.init_uniform(0.9)
    sim.moment_reset()
    sim.integrate_moment(1.0, 30)
    sim.integrate_moment(-1.0, 30)
    g = sim.get_state()
    mom, x0 = sim.get_moment()
    grads = np.array([[0.01, 0, 0], [0, 0.02, 0.01], [0, 0, 0]], np.float32)
    np.testing.assert_allclose(sim.signal_moment(grads, PC),
                               O.moment_signal(g["mag"], mom, g["comp"], grads, PC), rtol=1e-11,
                               atol=1e-14)


def test_mwc_stream_matches_oracle(ref_inputs):
    _say("case mwc")
    m = _mesh(ref_inputs, "sim.fibers_1_r004_f09", 6.0)
    n = 2000
    w, sub, sim = _pair(m, n=n, rng=D.RNG_MWC)
    pos, comp, fib = _spins(w, m, n, 5)
    seeds = np.random.default_rng(6).integers(1, 2 ** 32, (n, 2), dtype=np.uint64).astype(np.uint32)
    st = O.State(pos, comp=comp, fiber=fib, seed=seeds)
    O.integrate(w, st, 1, rng=O.RNG_MWC)
    sim.set_state(pos=pos, comp=comp, fiber=fib, seeds=seeds)
    sim.integrate(np.zeros(3, np.float32), 0.0, 1)
    g = sim.get_state(seeds=True)
    np.testing.assert_array_equal(g["seeds"], st.seed)      # integer stream is exact
    np.testing.assert_allclose(g["pos"], st.pos, rtol=1e-5, atol=2e-6)
    assert np.mean(g["comp"] == st.comp) > 0.99


def test_philox_stream_matches_oracle(ref_inputs):
    _say("case philox")
    m = _mesh(ref_inputs, "sim.fibers_1_r004_f09", 6.0)
    n = 4000
    dims, lo, cell = (31, 31, 7), (-1.1, -1.1, -1.1), (2.2 / 31, 2.2 / 31, 2.2 / 7)
    w, sub, sim = _pair(m, n=n, mode=1, grid="aabb", dims=dims, lo=lo, cell=cell, seed=0xABCDEF12345)
    pos, comp, fib = _spins(w, m, n, 5, plus_x=True)
    key = (0xABCDEF12345 & 0xFFFFFFFF, 0xABCDEF12345 >> 32)
    st = O.State(pos, comp=comp, fiber=fib, moment=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    sim.moment_reset()
    O.integrate(w, st, 20, rng=O.RNG_PHILOX, key=key, step0=0, weight=1.0)
    sim.integrate_moment(1.0, 20)
    g = sim.get_state()
    # hardware log2/cos units vs libm: normals agree to ~1e-6 relative
    close = np.all(np.abs(g["pos"] - st.pos) <= 1e-5 * np.maximum(1, np.abs(st.pos)), axis=1)
    assert close.mean() > 0.995
    assert np.mean(g["comp"] == st.comp) > 0.995


def test_free_diffusion_signal_matches_exp_minus_bD():
    """No membranes, PGSE 10/40/10 ms: S = exp(-b D) within 3 MC sigma."""
    _say("case free-diffusion")
    V = np.zeros((3, 3), np.float32) + 5
    n = 200000
    D_adc, space_scale = 2.0, 100.0
    sd = np.float32(np.sqrt(2 * D_adc * DT) / space_scale)
    w, sub, sim = _pair({"vertices": V, "trgls": [[0, 1, 2]], "tri_info": [[0, 0, 0]]}, n=n,
                        mode=1, grid="aabb", dims=(4, 4, 4), lo=(-1.1,) * 3, cell=(0.55,) * 3,
                        std=(sd,) * 4, t2=(np.inf,) * 4, perm=0.0, boundary=1)
    sim.init_uniform(1.0)
    sim.moment_reset()
    L, B = 2000, 8000
    sim.integrate_moment(1.0, L)
    sim.integrate_moment(0.0, B - L)
    sim.integrate_moment(-1.0, L)
    sim.integrate_moment(0.0, 2000)
    G = np.array([20.0, 40.0, 61.7])  # mT/m
    grads = np.zeros((3, 3), np.float32)
    grads[:, 0] = G * 1e-9 * space_scale
    S = sim.signal_moment(grads, PC)[:, 0]
    delta, Delta = 10.0, 40.0
    b = (2 * np.pi * GAMMA * G * 1e-9 * delta) ** 2 * (Delta - delta / 3)   # diffusion.cpp:317
    want = np.exp(-b * D_adc)
    sigma = np.sqrt((1 - want ** 2) / (2 * n)) + 1e-4
    assert np.all(np.abs(S - want) < 3 * sigma + 2e-3), (S, want)


def test_sorted_philox_run_is_order_independent(ref_inputs):
    """Spatial reordering and the launch length change only where spins live
    in memory and how the steps are grouped into launches: per-spin results
    (keyed by spin id) are bit-identical with and without them."""
    _say("case sort-invariance")
    m = _mesh(ref_inputs, "sim.fibers_1_r004_f09", 6.0)
    n = 20000
    dims, lo, cell = (41, 41, 9), (-1.2, -1.2, -1.2), (2.4 / 41, 2.4 / 41, 2.4 / 9)
    outs = []
    # (reorder interval, steps per launch): 100-step launches, 32-step
    # reorders, 7-step launches, 48-step reorders with long launches
    for sort, spl in ((0, 0), (32, 0), (0, 7), (48, 1024)):
        w, sub, sim = _pair(m, mode=1, grid="aabb", dims=dims, lo=lo, cell=cell, boundary=1, n=n,
                            sort_interval=sort, seed=77, steps_per_launch=spl)
        sim.init_uniform(1.0)
        sim.label(m["xmax"], m["ymax"], m["zmax"])
        sim.moment_reset()
        sim.integrate_moment(1.0, 100)
        sim.integrate_moment(-1.0, 100)
        st = sim.get_state()
        mom, x0 = sim.get_moment()
        sig = sim.signal_moment(np.array([[0.02, 0.01, 0.0]], np.float32), PC)
        outs.append((st, mom, x0, sig))
        sim.close()
    a, ma, xa, sa = outs[0]
    for b, mb, xb, sb in outs[1:]:
        for k in ("pos", "mag", "comp", "fiber"):
            np.testing.assert_array_equal(a[k], b[k])
        np.testing.assert_array_equal(ma, mb)
        np.testing.assert_array_equal(xa, xb)
        np.testing.assert_allclose(sa, sb, rtol=1e-12)


def test_fresh_mwc_handle_walks_a_random_stream():
    # a handle created with MWC draws and no seeds of its own starts from
    # distinct nonzero seed pairs (an all-zero multiply-with-carry state would
    # repeat one draw every step): the free walk has zero mean and variance
    # iters * sd^2 per axis
    _say("case fresh mwc seeds")
    V = np.zeros((3, 3), np.float32) + 5
    n, iters, sd = 20000, 100, 0.01
    w, sub, sim = _pair({"vertices": V, "trgls": [[0, 1, 2]], "tri_info": [[0, 0, 0]]}, n=n,
                        rng=D.RNG_MWC, std=(sd,) * 4, seed=9)
    sim.init_uniform(0.5)
    g0 = sim.get_state(seeds=True)
    seeds = g0["seeds"]
    assert (seeds != 0).all()
    assert len(np.unique(seeds.view(np.uint64))) == n
    sim.integrate(np.zeros(3, np.float32), 0.0, iters)
    d = (sim.get_state()["pos"] - g0["pos"]).astype(np.float64)
    se = sd * np.sqrt(iters / n)
    assert np.all(np.abs(d.mean(0)) < 5 * se), d.mean(0)
    np.testing.assert_allclose(d.var(0), iters * sd * sd, rtol=0.05)


def test_reference_mode_stops_a_step_that_never_ends():
    # the step of the small configuration on which the reference's collision
    # loop never ends (tests/test_oracle_pins.py): the kernel stops the spin
    # after 4096 collisions, the call reports DSIM_ERR_OVERFLOW once, the
    # state stays valid and the walk continues; in the physical mode the same
    # step ends and matches the oracle bit for bit
    _say("case never-ending reference step")
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_oracle_pins import _shared_diagonal_case
    env, pos, q = _shared_diagonal_case()
    V, T, I = env["vertices"], env["triangles"], env["tri_info"]
    sub = D.Substrate(V, T, I)
    g0 = np.zeros(3, np.float32)
    sim = D.Simulation(sub, 1, dt=DT, std_dev=env["std_dev"], t2=env["t2"], mode=D.MODE_REFERENCE,
                       rng=D.RNG_PHILOX)
    sim.set_state(pos=pos, comp=np.zeros(1, np.uint8), fiber=np.full(1, 65535, np.uint16))
    with pytest.raises(D.DsimError) as err:
        sim.integrate_injected(g0, 0.0, q, trace=True)
    assert err.value.code == -5                  # DSIM_ERR_OVERFLOW
    g = sim.get_state()                          # the state is readable again
    assert np.all(np.abs(g["pos"]) <= 1.0)
    q2 = np.array([[[0.1, 0.2, 0.3, 0.5]]], np.float32)
    sim.integrate_injected(g0, 0.0, q2)          # and the walk goes on
    sim.sync()
    sim.close()
    w = O.World(V, T, I, std_dev=env["std_dev"], t2=env["t2"], dt=DT, permeability=0.0)
    w.set(test=1, traversal=1, transitions=1)
    st = O.State(pos, comp=np.zeros(1, np.uint8), fiber=np.full(1, 65535, np.uint16))
    tr_o = O.integrate(w, st, 1, rng=O.RNG_INJECT, inject=q, trace=True)
    dims, lo, cell = (50, 50, 50), (-1.0, -1.0, -1.0), (0.04, 0.04, 0.04)
    sub2 = D.Substrate(V, T, I, grid="aabb", dims=dims, lo=lo, cell=cell, eps=1e-5)
    w2 = O.World(V, T, I, std_dev=env["std_dev"], t2=env["t2"], dt=DT, permeability=0.0,
                 grid="aabb", dims=dims, lo=lo, cell=cell, eps=1e-5)
    w2.set(test=1, traversal=1, transitions=1)
    st2 = O.State(pos, comp=np.zeros(1, np.uint8), fiber=np.full(1, 65535, np.uint16))
    tr_o2 = O.integrate(w2, st2, 1, rng=O.RNG_INJECT, inject=q, trace=True)
    sim = D.Simulation(sub2, 1, dt=DT, std_dev=env["std_dev"], t2=env["t2"],
                       mode=D.MODE_PHYSICAL, rng=D.RNG_PHILOX)
    sim.set_state(pos=pos, comp=np.zeros(1, np.uint8), fiber=np.full(1, 65535, np.uint16))
    tr_g = sim.integrate_injected(g0, 0.0, q, trace=True)
    assert tr_o[0, 0, 2] < 10 and tr_o2[0, 0, 2] < 10
    np.testing.assert_array_equal(tr_g[..., [0, 1, 2, 4]], tr_o2[..., [0, 1, 2, 4]])
    np.testing.assert_array_equal(sim.get_state()["pos"], st2.pos)
    sim.close()

"""The small benchmark configuration end to end on the GPU against the oracle:
1e4 spins, 16 capped 32-gon axons in a 20 um periodic voxel, PGSE 10/40 ms
over 12,000 steps, 6 directions at b = 1 ms/um^2 and b = 0.  Both walk the same
Philox streams from the same starts; the hardware log2/sin/cos units differ
from libm in the last bits, so a few trajectories part ways at a wall and the
signals are compared within 3 Monte Carlo sigma."""
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D
from dsim_amd import substrate as S

pytestmark = pytest.mark.gpu

GAMMA = 42576.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def test_small_configuration_free_run_matches_oracle():
    _say("case small configuration, 12000 steps")
    cfg = S.SMALL
    env = S.build(cfg, cell_um=0.5, axons=S.pack(cfg))
    g = env["grid"]
    n, seed = cfg.n_spins, 0x5EED00000001
    sub = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], grid="aabb",
                      dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"], device=True)
    sim = D.Simulation(sub, n, dt=cfg.dt, std_dev=env["std_dev"], t2=env["t2"],
                       permeability=env["permeability"], mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=seed)
    w = O.World(env["vertices"], env["triangles"], env["tri_info"], grid="aabb", dims=g["dims"],
                lo=g["lo"], cell=g["cell"], eps=g["eps"], std_dev=env["std_dev"], t2=env["t2"],
                dt=cfg.dt, permeability=env["permeability"], test=1, traversal=1,
                transitions=1, boundary=1)
    pos = np.random.default_rng(31).uniform(-1, 1, (n, 3)).astype(np.float32)
    comp, fib, _, _ = w.find_spins(pos, env["xmax"], env["xmax"], env["xmax"], plus_x=True)
    st = O.State(pos, comp=comp, fiber=fib, moment=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    sim.moment_reset()
    key = (seed & 0xFFFFFFFF, seed >> 32)
    L, B, TE = (int(round(v / cfg.dt)) for v in (cfg.delta, cfg.Delta, cfg.TE))
    step = 0
    for wgt, k in ((1.0, L), (0.0, B - L), (-1.0, L), (0.0, TE - B - L)):
        sim.integrate_moment(wgt, k)
        O.integrate(w, st, k, rng=O.RNG_PHILOX, key=key, step0=step, weight=wgt)
        step += k
    assert step == 12000
    prot = env["protocol"]
    grads = (prot[:, 3:6].astype(np.float64) * 1e-9 * env["space_scale"]).astype(np.float32)
    pc = np.float32(2 * np.pi * GAMMA * cfg.dt)
    got = sim.signal_moment(grads, pc)
    want = O.moment_signal(st.mag, st.moment, st.comp, grads, pc)
    gs = sim.get_state()
    same = np.all(np.abs(gs["pos"] - st.pos) <= 1e-4 * np.maximum(1, np.abs(st.pos)), axis=1)
    sim.close()
    _say(f"  {same.mean():.4f} of spins end within 1e-4; total signal GPU "
         f"{np.round(got[:, 0], 5)} oracle {np.round(want[:, 0], 5)}")
    assert same.mean() > 0.9
    counts = np.bincount(st.comp, minlength=4)
    for c in range(5):
        m = n if c == 0 else counts[c - 1]
        if m < 100:
            continue
        sigma = np.sqrt(np.maximum(1 - want[:, c] ** 2, 1e-6) / (2 * m))
        assert np.all(np.abs(got[:, c] - want[:, c]) < 3 * sigma), (c, got[:, c], want[:, c])


def test_scale_configuration_sample_matches_oracle():
    # the at-scale substrate (2000 packed axons, 64-gon open prisms, 0.2 um
    # cells) on a 2e4-spin sample, with the spin reordering of the benchmark
    _say("case scale configuration sample, 1500 steps")
    cfg = S.SCALE
    env = S.build(cfg, cell_um=0.2, axons=S.pack(cfg))
    g = env["grid"]
    n, seed = 20000, 0x5EED00000002
    sub = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], grid="aabb",
                      dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"], device=True)
    sim = D.Simulation(sub, n, dt=cfg.dt, std_dev=env["std_dev"], t2=env["t2"],
                       permeability=env["permeability"], mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=seed,
                       sort_interval=256)
    w = O.World(env["vertices"], env["triangles"], env["tri_info"], grid="aabb", dims=g["dims"],
                lo=g["lo"], cell=g["cell"], eps=g["eps"], std_dev=env["std_dev"], t2=env["t2"],
                dt=cfg.dt, permeability=env["permeability"], test=1, traversal=1,
                transitions=1, boundary=1)
    pos = np.random.default_rng(32).uniform(-1, 1, (n, 3)).astype(np.float32)
    comp, fib, _, _ = w.find_spins(pos, env["xmax"], env["xmax"], env["xmax"], plus_x=True)
    st = O.State(pos, comp=comp, fiber=fib, moment=True)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    sim.moment_reset()
    key = (seed & 0xFFFFFFFF, seed >> 32)
    step = 0
    for wgt, k in ((1.0, 500), (0.0, 500), (-1.0, 500)):
        sim.integrate_moment(wgt, k)
        O.integrate(w, st, k, rng=O.RNG_PHILOX, key=key, step0=step, weight=wgt)
        step += k
    # delta = 2.5 ms, Delta = 5 ms: gradients for b = 0.5 and 2 ms/um^2 along x, y, z
    q = np.sqrt(np.array([0.5, 2.0]) / (5.0 - 2.5 / 3)) / 2.5    # gamma G in 1/(um ms)
    G = q / (2 * np.pi * GAMMA * 1e-9)                           # mT/m
    grads = np.zeros((7, 3), np.float32)
    for j, gg in enumerate(G):
        for a in range(3):
            grads[1 + 3 * j + a, a] = gg * 1e-9 * env["space_scale"]
    pc = np.float32(2 * np.pi * GAMMA * cfg.dt)
    got = sim.signal_moment(grads, pc)
    want = O.moment_signal(st.mag, st.moment, st.comp, grads, pc)
    gs = sim.get_state()
    same = np.all(np.abs(gs["pos"] - st.pos) <= 1e-4 * np.maximum(1, np.abs(st.pos)), axis=1)
    sim.close()
    _say(f"  {same.mean():.4f} of spins end within 1e-4; total signal GPU "
         f"{np.round(got[:, 0], 5)} oracle {np.round(want[:, 0], 5)}")
    assert same.mean() > 0.9
    assert np.array_equal(gs["comp"], st.comp) or np.mean(gs["comp"] == st.comp) > 0.99
    counts = np.bincount(st.comp, minlength=4)
    for c in range(5):
        m = n if c == 0 else counts[c - 1]
        if m < 100:
            continue
        sigma = np.sqrt(np.maximum(1 - want[:, c] ** 2, 1e-6) / (2 * m))
        assert np.all(np.abs(got[:, c] - want[:, c]) < 3 * sigma), (c, got[:, c], want[:, c])

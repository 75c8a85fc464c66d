"""The per-step phase path on the GPU (dsim_integrate, the twin of
integrateSystem, spinSystem.cu:221-270): PGSE lobes of +G and -G through the
reference-mode kernel and through the physical-mode walk in a periodic
voxel (cubic and not), against the oracle on the same Philox streams and
against the exact attenuation of a free Gaussian walk; and the device spin
draw of resetSpins (spinSystem.cpp:1386-1407)."""
import sys

import numpy as np
import pytest

import oracle as O
import pgse
from dsim_amd import lib as D

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


FAR = (np.array([[5.0, 5.0, 5.0], [5.001, 5.0, 5.0], [5.0, 5.001, 5.0]], np.float32),
       np.array([[0, 1, 2]], np.uint32), np.array([[0, 0, 0]], np.uint32))
KEY = (1, 2)
SEED = KEY[0] | (KEY[1] << 32)


def _run_pgse(sim, world, n, G_um, space, dt, start, gdir=(0.6, 0.0, 0.8)):
    sched, _ = O.pgse_schedule(2.0, 6.0, 0.0, dt, 10)
    gs = np.float32(G_um * 1e-9 * space)
    g = np.asarray(gdir, np.float32) * gs
    pc = np.float32(2 * np.pi * pgse.GAMMA * dt)
    rng = np.random.default_rng(31)
    pos = rng.uniform(-start, start, (n, 3)).astype(np.float32)
    st = O.State(pos)
    sim.set_state(pos=pos, comp=np.zeros(n, np.uint8))
    step = [0]

    def adv(sign, steps):
        gg = (g * sign).astype(np.float32)
        sim.integrate(gg, pc, steps)
        O.integrate(world, st, steps, rng=O.RNG_PHILOX, key=KEY, step0=step[0], grad=gg,
                    phase_constant=pc)
        step[0] += steps

    pgse.run_updates(sched, 10, adv)
    sig_gpu = sim.signal()[0]
    sig_orc = O.mr_signal(st.mag, st.phase, st.comp)[0]
    sd = world.params["std_dev"][0]
    want, v = pgse.walk_attenuation(pgse.step_signs(sched, 10), gs, pc, sd)
    return sig_gpu, sig_orc, want, pgse.mc_sigma(v, n), sim.get_state(), st


def test_pgse_reference_mode_phase_path():
    space, D_, dt = 100.0, 2.0, 0.005
    sd = np.float32(np.sqrt(2 * D_ * dt) / space)
    n = 20000
    _say(f"reference-mode PGSE: {n} spins, 1600 steps")
    w = O.World(*FAR, std_dev=(sd,) * 4, dt=dt)
    sub = D.Substrate(*FAR)
    sim = D.Simulation(sub, n, dt=dt, std_dev=(sd,) * 4, mode=D.MODE_REFERENCE,
                       rng=D.RNG_PHILOX, seed=SEED)
    g, o, want, sigma, gs, st = _run_pgse(sim, w, n, 600.0, space, dt, start=0.2)
    _say(f"  gpu {g:.6f} oracle {o:.6f} gaussian {want:.6f} sigma {sigma:.2e}")
    # same streams: the positions agree to the last bits of the hardware
    # transcendental units, the signals far inside the MC error
    np.testing.assert_allclose(gs["pos"], st.pos, atol=1e-5)
    assert abs(g - o) < 1e-3
    assert abs(g - want) < 3 * sigma
    assert abs(o - want) < 3 * sigma


@pytest.mark.parametrize("voxel", [None, ((-0.5, -1.0, -0.25), (0.5, 1.0, 0.25))])
def test_pgse_periodic_voxel_phase_path(voxel):
    # most spins cross a face during the sequence; the phase uses the
    # unwrapped position on both sides
    space, D_, dt = 5.0, 2.0, 0.005
    sd = np.float32(np.sqrt(2 * D_ * dt) / space)
    n = 20000
    _say(f"periodic PGSE, voxel {voxel}: {n} spins, 1600 steps")
    kw = {} if voxel is None else dict(voxel_lo=voxel[0], voxel_hi=voxel[1])
    grid = dict(grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1), cell=(2, 2, 2), eps=0.0)
    w = O.World(*FAR, std_dev=(sd,) * 4, dt=dt, boundary=1, test=1, traversal=1,
                transitions=1, **grid, **kw)
    sub = D.Substrate(*FAR, **grid)
    sim = D.Simulation(sub, n, dt=dt, std_dev=(sd,) * 4, mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=SEED, voxel=voxel)
    g, o, want, sigma, gs, st = _run_pgse(sim, w, n, 600.0, space, dt, start=0.0)
    _say(f"  gpu {g:.6f} oracle {o:.6f} gaussian {want:.6f} sigma {sigma:.2e}")
    lo = np.array(voxel[0] if voxel else (-1, -1, -1), np.float32)
    hi = np.array(voxel[1] if voxel else (1, 1, 1), np.float32)
    assert np.all(gs["pos"] >= lo) and np.all(gs["pos"] < hi)
    np.testing.assert_allclose(gs["pos"], st.pos, atol=1e-5)
    assert abs(g - o) < 1e-3
    assert abs(g - want) < 3 * sigma
    assert abs(o - want) < 3 * sigma


def test_periodic_injected_phase_bit_exact():
    # injected normals: positions, crossings and phases equal the oracle's
    # bit for bit in a non-cubic periodic voxel
    n, iters = 2048, 300
    voxel = ((-0.3, -1.0, -0.2), (0.3, 1.0, 0.2))
    sd = np.float32(0.05)
    grid = dict(grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1), cell=(2, 2, 2), eps=0.0)
    w = O.World(*FAR, std_dev=(sd,) * 4, boundary=1, test=1, traversal=1, transitions=1,
                voxel_lo=voxel[0], voxel_hi=voxel[1], **grid)
    sub = D.Substrate(*FAR, **grid)
    sim = D.Simulation(sub, n, dt=0.005, std_dev=(sd,) * 4, mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, voxel=voxel)
    rng = np.random.default_rng(3)
    pos = rng.uniform(voxel[0], voxel[1], (n, 3)).astype(np.float32)
    q = np.empty((iters, n, 4), np.float32)
    q[..., :3] = rng.standard_normal((iters, n, 3))
    q[..., 3] = 0.5
    g = np.array([3e-3, -2e-3, 5e-3], np.float32)
    pc = np.float32(1337.0)
    st = O.State(pos)
    O.integrate(w, st, iters, rng=O.RNG_INJECT, inject=q, grad=g, phase_constant=pc)
    sim.set_state(pos=pos, comp=np.zeros(n, np.uint8))
    sim.integrate_injected(g, pc, q)
    gs = sim.get_state()
    assert np.abs(st.wrap).max() >= 3  # spins crossed faces several times
    assert np.array_equal(gs["pos"], st.pos)
    assert np.array_equal(gs["phase"], st.phase)


def test_init_spins_uniform_pin():
    # resetSpins (spinSystem.cpp:1386-1398): positions uniform in the box,
    # magnitude 1, phase 0, labels cleared
    from scipy import stats
    n = 200000
    sub = D.Substrate(*FAR)
    sim = D.Simulation(sub, n, dt=0.005, std_dev=(0.01,) * 4, seed=77)
    sim.set_state(pos=np.zeros((n, 3), np.float32), comp=np.full(n, 3, np.uint8),
                  fiber=np.full(n, 7, np.uint16), mag=np.full(n, 0.5, np.float32),
                  phase=np.full(n, 2.0, np.float32))
    b = 0.75
    sim.init_uniform(b)
    s = sim.get_state()
    p = s["pos"]
    assert np.all(p >= -b) and np.all(p <= b)
    for k in range(3):
        assert stats.kstest((p[:, k] + b) / (2 * b), "uniform").pvalue > 1e-3
    assert abs(np.corrcoef(p.T)[0, 1]) < 0.01 and abs(np.corrcoef(p.T)[1, 2]) < 0.01
    assert np.all(s["mag"] == 1.0) and np.all(s["phase"] == 0.0)
    assert np.all(s["comp"] == 0) and np.all(s["fiber"] == D.NO_FIBER)
    # the box form draws the same stream: the cube gives the same bits,
    # another box maps it affinely per axis
    sim.init_box((-b, -b, -b), (b, b, b))
    assert np.array_equal(sim.get_state()["pos"], p)
    lo, hi = np.array([-0.2, -1.0, 0.1], np.float32), np.array([0.4, 1.0, 0.3], np.float32)
    sim.init_box(lo, hi)
    q = sim.get_state()["pos"]
    assert np.all(q >= lo) and np.all(q <= hi)
    u = (p + b) / (2 * b)
    np.testing.assert_allclose(q, lo + (hi - lo) * u, atol=2e-7)
    # a different seed draws different positions
    sim2 = D.Simulation(sub, n, dt=0.005, std_dev=(0.01,) * 4, seed=78)
    sim2.init_uniform(b)
    assert np.mean(sim2.get_state()["pos"] == p) < 1e-3

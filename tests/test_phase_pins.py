"""Pins of the oracle's per-step phase (spinKernelCpu.cpp:490, oracle
step_one) that do not retype its formula: a one-step known answer fixes the
post-step position and the axis pairing, and PGSE walks against the exact
Gaussian attenuation fix the lobe signs, the phase constant and, in the
periodic voxel, the unwrapping of positions that cross a face."""
import numpy as np
import pytest

import oracle as O
import pgse


def _far_world(**kw):
    """A world whose one (tiny) triangle lies outside every path of the test."""
    V = np.array([[5.0, 5.0, 5.0], [5.001, 5.0, 5.0], [5.0, 5.001, 5.0]], np.float32)
    T = np.array([[0, 1, 2]], np.uint32)
    I = np.array([[0, 0, 0]], np.uint32)
    return O.World(V, T, I, **kw)


def test_phase_known_answer_post_step_position():
    # one spin, two injected steps: the phase after step t adds pc * g . x_t
    # with x_t the position AFTER the step (a pre-step rule gives 0 for the
    # first step), each axis of g paired with the same axis of x
    sd = 0.125
    w = _far_world(std_dev=(sd,) * 4)
    st = O.State(np.zeros((1, 3), np.float32))
    q = np.array([[[1.0, 2.0, -4.0, 0.5]], [[2.0, 0.0, 1.0, 0.5]]], np.float32)
    g = np.array([0.5, 0.25, 2.0], np.float32)
    pc = 2.0
    O.integrate(w, st, 1, rng=O.RNG_INJECT, inject=q[:1], grad=g, phase_constant=pc)
    x1 = np.array([0.125, 0.25, -0.5])
    assert st.phase[0] == pytest.approx(pc * float(g @ x1), abs=0)  # exact in binary
    assert st.phase[0] == -1.75
    O.integrate(w, st, 1, rng=O.RNG_INJECT, inject=q[1:], grad=g, phase_constant=pc)
    x2 = x1 + sd * np.array([2.0, 0.0, 1.0])
    assert st.phase[0] == pytest.approx(pc * float(g @ x1) + pc * float(g @ x2), abs=0)


def _pgse(world, n, sd, G_um, delta, Delta, dt, s, rng, *, start, wrap=True, seed_key=(1, 2)):
    """Runs PGSE through orc_integrate's per-step phase; returns (signal,
    expected walk attenuation, MC sigma, exp(-bD))."""
    sched, (L, B, R, _) = O.pgse_schedule(delta, Delta, 0.0, dt, s)
    gs = np.float32(G_um * 1e-9 * world.space_scale)
    gdir = np.array([0.6, 0.0, 0.8], np.float32)
    pc = np.float32(2 * np.pi * pgse.GAMMA * dt)
    pos = rng.uniform(-start, start, (n, 3)).astype(np.float32)
    st = O.State(pos, wrap=wrap)
    step = [0]

    def adv(sign, steps):
        O.integrate(world, st, steps, rng=O.RNG_PHILOX, key=seed_key, step0=step[0],
                    grad=(gdir * gs * sign).astype(np.float32), phase_constant=pc)
        step[0] += steps

    pgse.run_updates(sched, s, adv)
    sig = O.mr_signal(st.mag, st.phase, st.comp)[0]
    want, v = pgse.walk_attenuation(pgse.step_signs(sched, s), gs, pc, sd)
    D = (sd * world.space_scale) ** 2 / (2 * dt)
    return sig, want, pgse.mc_sigma(v, n), np.exp(-pgse.b_value(G_um, delta, Delta) * D)


def test_walk_attenuation_tends_to_exp_bD():
    # the exact discrete attenuation against the reference's b (diffusion.cpp:317)
    dt, s, delta, Delta, G = 0.005, 10, 2.0, 6.0, 600.0
    sched, _ = O.pgse_schedule(delta, Delta, 0.0, dt, s)
    space, D = 100.0, 2.0
    sd = np.sqrt(2 * D * dt) / space
    gs = G * 1e-9 * space
    pc = 2 * np.pi * pgse.GAMMA * dt
    want, _ = pgse.walk_attenuation(pgse.step_signs(sched, s), gs, pc, sd)
    bD = pgse.b_value(G, delta, Delta) * D
    assert 0.8 < bD < 3
    assert want == pytest.approx(np.exp(-bD), rel=0.01)


def test_pgse_reference_mode_clamp_box_matches_gaussian():
    # reference semantics (softsurfer test, first-cell traversal, clamp at
    # 1 - 0.005); spins start near the centre and never reach the clamp
    space, D, dt = 100.0, 2.0, 0.005
    sd = np.float32(np.sqrt(2 * D * dt) / space)
    w = _far_world(std_dev=(sd,) * 4, dt=dt)
    w.space_scale = space
    rng = np.random.default_rng(11)
    n = 12000
    sig, want, sigma, ebd = _pgse(w, n, sd, 600.0, 2.0, 6.0, dt, 10, rng, start=0.2)
    assert abs(sig - want) < 3 * sigma, (sig, want, sigma)
    assert abs(sig - ebd) < 3 * sigma + 0.01 * ebd
    assert 0.1 < want < 0.6


def test_pgse_wrong_second_lobe_sign_is_detected():
    # a planted mistake in the schedule (both lobes +G) must fail the check
    space, D, dt = 100.0, 2.0, 0.005
    sd = np.float32(np.sqrt(2 * D * dt) / space)
    w = _far_world(std_dev=(sd,) * 4, dt=dt)
    rng = np.random.default_rng(12)
    n = 4000
    sched, _ = O.pgse_schedule(2.0, 6.0, 0.0, dt, 10)
    bad = np.abs(sched)
    gs = np.float32(600.0 * 1e-9 * space)
    pc = np.float32(2 * np.pi * pgse.GAMMA * dt)
    st = O.State(rng.uniform(-0.2, 0.2, (n, 3)).astype(np.float32))
    g = np.array([0.6, 0.0, 0.8], np.float32) * gs
    step = [0]

    def adv(sign, steps):
        O.integrate(w, st, steps, rng=O.RNG_PHILOX, key=(1, 1), step0=step[0],
                    grad=(g * sign).astype(np.float32), phase_constant=pc)
        step[0] += steps

    pgse.run_updates(bad, 10, adv)
    sig = O.mr_signal(st.mag, st.phase, st.comp)[0]
    want, v = pgse.walk_attenuation(pgse.step_signs(sched, 10), gs, pc, sd)
    assert abs(sig - want) > 10 * pgse.mc_sigma(v, n)


@pytest.mark.parametrize("voxel", [None, ((-0.5, -1.0, -0.25), (0.5, 1.0, 0.25))])
def test_pgse_periodic_voxel_unwraps_the_phase(voxel):
    # physical mode in a small periodic voxel: most spins cross a face during
    # the sequence; the phase must use the unwrapped position
    space, D, dt = 5.0, 2.0, 0.005
    sd = np.float32(np.sqrt(2 * D * dt) / space)
    kw = {} if voxel is None else dict(voxel_lo=voxel[0], voxel_hi=voxel[1])
    w = _far_world(std_dev=(sd,) * 4, dt=dt, grid="aabb", dims=(1, 1, 1), lo=(-1, -1, -1),
                   cell=(2, 2, 2), eps=0.0, boundary=1, test=1, traversal=1, transitions=1, **kw)
    w.space_scale = space
    rng = np.random.default_rng(13)
    n = 12000
    G = 600.0
    sig, want, sigma, ebd = _pgse(w, n, sd, G, 2.0, 6.0, dt, 10, rng, start=0.0)
    # (start=0: spins begin at the centre; the voxel is small enough that
    # the walk crosses its faces)
    assert abs(sig - want) < 3 * sigma, (sig, want, sigma)
    # the same walk with the phase of the wrapped position is far off
    rng = np.random.default_rng(13)
    sig_w, _, _, _ = _pgse(w, n, sd, G, 2.0, 6.0, dt, 10, rng, start=0.0, wrap=False)
    assert abs(sig_w - want) > 10 * sigma, (sig_w, want)

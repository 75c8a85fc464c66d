"""Waveform phases on the matrix cores (include/dsim_phase.h) against FP64:
the accumulated phase of every (spin, waveform) pair against an FP64 product
of the recorded trajectory and the gradients, the signals against FP64 sums,
and a PGSE-shaped waveform against the gradient-moment path."""
import numpy as np
import pytest

from dsim_amd import lib as D
from dsim_amd import phase as PH
from dsim_amd import substrate as S

pytestmark = pytest.mark.gpu

GAMMA = 42576.0


@pytest.fixture(scope="module")
def env():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    cfg = S.SMALL
    e = S.build(cfg, cell_um=0.5, axons=S.pack(cfg))
    g = e["grid"]
    e["sub"] = D.Substrate(e["vertices"], e["triangles"], e["tri_info"], grid="aabb",
                           dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"])
    return e


def _sim(env, n, seed, sort_interval=0, rng=D.RNG_PHILOX):
    cfg = env["cfg"]
    sim = D.Simulation(env["sub"], n, dt=cfg.dt, std_dev=env["std_dev"],
                       t2=(90.0, 70.0, 70.0, 70.0), permeability=0.02, mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, rng=rng, seed=seed,
                       sort_interval=sort_interval)
    sim.init_uniform(1.0)
    sim.label(env["xmax"], env["xmax"], env["xmax"])
    sim.moment_reset()
    return sim


def _fp64_signal(phase, mag, comp):
    out = np.zeros((phase.shape[1], 5))
    z = mag[:, None] * np.exp(1j * phase)
    out[:, 0] = np.abs(z.sum(0)) / len(mag)
    for c in range(4):
        sel = comp == c
        if sel.any():
            out[:, 1 + c] = np.abs(z[sel].sum(0)) / sel.sum()
    return out


# 256 waveforms take the 128 x 256 workgroup tile, 384 (three tiles of 128) the
# 128 x 128 one, 200 are padded with zero waveforms to 256; 3000 spins leave a
# partly filled last spin tile
@pytest.mark.parametrize("n_wave", [256, 384, 200])
def test_phase_gemm_matches_fp64(env, n_wave):
    n, steps, chunk = 3000, 256, 64
    pc = np.float32(2 * np.pi * GAMMA * env["cfg"].dt)
    # 20x the 80 mT/m bound so that the phases reach a sizeable fraction of a radian
    g = PH.random_waveforms(n_wave, steps, seed=11, g_max=1600.0) * np.float32(
        1e-9 * env["space_scale"])
    sim = _sim(env, n, seed=21)
    wf = PH.Waveforms(sim, g, pc, chunk_steps=chunk)
    traj = wf.run(steps, trajectory=True)            # [steps, 3, n]
    phi, comp = wf.phase()
    # FP64 product of the same trajectory and the same fp32 coefficients
    coef = (pc * g).astype(np.float32).astype(np.float64)      # [steps, n_wave, 3]
    A = traj.transpose(2, 0, 1).reshape(n, 3 * steps).astype(np.float64)
    B = coef.transpose(0, 2, 1).reshape(3 * steps, n_wave)
    ref = A @ B
    scale = np.abs(A) @ np.abs(B)
    err = np.abs(phi.astype(np.float64) - ref)
    # bf16 hi/lo products carry ~2^-16 relative error each; FP32 sums per chunk
    # each product carries at most ~3 * 2^-18 relative error from the split and
    # the FP32 sums of a chunk about K * 2^-24; a hi-only product would be off
    # by ~2^-9
    assert (err <= 5e-5 * scale + 1e-7).all(), float((err / (scale + 1e-30)).max())
    assert np.abs(ref).max() > 0.1           # the test exercises non-trivial phases
    # the Kahan compensations are values of bf16 (low 16 bits zero) within an
    # ulp of their own phase (the last fold's rounding error); a pair word
    # decoded into the wrong spin would break the bound wherever neighbouring
    # phases differ in magnitude
    cb = comp.view(np.uint32)
    assert not (cb & 0xFFFF).any()
    ulp = np.spacing(np.abs(phi)).astype(np.float64)
    assert (np.abs(comp.astype(np.float64)) <= ulp * 1.01).all()
    assert np.count_nonzero(comp) > comp.size // 10   # the compensation is in use
    # signals: FP64 readout of the same phases (x0 term included)
    st = sim.get_state()
    _, x0 = sim.get_moment()
    total = ref + x0.astype(np.float64) @ coef.sum(0).T
    want = _fp64_signal(total, st["mag"].astype(np.float64), st["comp"])
    got = wf.signal()
    assert want[:, 0].min() < 0.99
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)
    wf.close()
    sim.close()


# Philox walks record trajectories four steps at a time (k_walk_traj); MWC
# walks store every step (the walk's general path)
@pytest.mark.parametrize("rng_kind", [D.RNG_PHILOX, D.RNG_MWC])
def test_pgse_waveform_matches_moment_path(env, rng_kind):
    n, steps, chunk = 4000, 320, 64
    L, B = 64, 192
    pc = np.float32(2 * np.pi * GAMMA * env["cfg"].dt)
    rng = np.random.default_rng(5)
    dirs = rng.normal(size=(256, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    G = (dirs * 3000.0 * 1e-9 * env["space_scale"]).astype(np.float32)   # [n_wave, 3]
    w = np.zeros(steps, np.float32)
    w[:L] = 1.0
    w[B:B + L] = -1.0
    g = w[:, None, None] * G[None, :, :]
    sim = _sim(env, n, seed=33, rng=rng_kind)
    wf = PH.Waveforms(sim, g, pc, chunk_steps=chunk)
    wf.run(steps)
    got = wf.signal()
    wf.close()
    sim.close()
    ref_sim = _sim(env, n, seed=33, rng=rng_kind)
    for wt, k in [(1.0, L), (0.0, B - L), (-1.0, L), (0.0, steps - B - L)]:
        ref_sim.integrate_moment(wt, k)
    want = ref_sim.signal_moment(G, pc)
    ref_sim.close()
    assert want[:, 0].min() < 0.99     # the gradients attenuate
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)


def test_trajectory_buffer_changes_nothing(env):
    # the Kahan folds fall every chunk whatever the buffer: one launch over a
    # buffer of several chunks gives the phases of one launch per chunk, bit
    # for bit (here 320 steps: a buffer of 256 steps, then one of 64)
    n, n_wave, steps, chunk = 3000, 256, 320, 64
    pc = np.float32(2 * np.pi * GAMMA * env["cfg"].dt)
    g = PH.random_waveforms(n_wave, steps, seed=12, g_max=1600.0) * np.float32(
        1e-9 * env["space_scale"])
    out = []
    for buffer in (64, 256):
        sim = _sim(env, n, seed=22)
        wf = PH.Waveforms(sim, g, pc, chunk_steps=chunk, buffer_steps=buffer)
        wf.run(steps)
        out.append(wf.phase() + (wf.signal(),))
        wf.close()
        sim.close()
    (p0, c0, s0), (p1, c1, s1) = out
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_array_equal(s0, s1)
    assert np.abs(p0).max() > 0.1


def test_reordered_spins_give_the_same_phases(env):
    # with spatial reordering between walk launches the trajectory rows and
    # phases follow the spin id: the same phases bit for bit, and the signals
    # (FP64 sums over slots, in another order) to rounding
    n, n_wave, steps, chunk = 3000, 256, 512, 64
    pc = np.float32(2 * np.pi * GAMMA * env["cfg"].dt)
    g = PH.random_waveforms(n_wave, steps, seed=13, g_max=1600.0) * np.float32(
        1e-9 * env["space_scale"])
    out = []
    for sort in (0, 128):
        sim = _sim(env, n, seed=24, sort_interval=sort)
        wf = PH.Waveforms(sim, g, pc, chunk_steps=chunk, buffer_steps=256)
        wf.run(steps)
        out.append(wf.phase() + (wf.signal(),))
        wf.close()
        sim.close()
    (p0, c0, s0), (p1, c1, s1) = out
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_allclose(s0, s1, rtol=0, atol=1e-12)
    assert np.abs(p0).max() > 0.1

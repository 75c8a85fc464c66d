"""Free-run physics on the GPU against closed forms: PGSE signal of spins
inside an impermeable cylinder (a 64-gon prism through the periodic voxel)
across the axis against the Gaussian phase approximation, and along the axis
against free diffusion, each within 3 Monte Carlo sigma plus the stated model
allowance."""
import sys

import numpy as np
import pytest

import analytic as A
from dsim_amd import lib as D

pytestmark = pytest.mark.gpu

DT = 0.005                    # ms
D_ADC = 2.0                   # um^2/ms
SPACE = 10.0                  # um per normalised unit: the voxel [-1, 1)^3 is 20 um
R_UM = 4.0
NGON = 64
GAMMA = 42576.0
PC = np.float32(2 * np.pi * GAMMA * DT)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _prism(r, z0, z1, ngon):
    ang = 2 * np.pi * np.arange(ngon) / ngon
    ring = np.stack([r * np.cos(ang), r * np.sin(ang)], 1)
    V = np.concatenate([np.c_[ring, np.full(ngon, z0)], np.c_[ring, np.full(ngon, z1)]])
    b = np.arange(ngon)
    t = b + ngon
    bn, tn = np.roll(b, -1), np.roll(t, -1)
    T = np.concatenate([np.stack([b, bn, tn], 1), np.stack([b, tn, t], 1)])
    info = np.zeros_like(T)
    info[:, 2] = T[:, 0]
    return V.astype(np.float32), T.astype(np.uint32), info.astype(np.uint32)


def test_cylinder_signal_matches_gaussian_phase_approximation():
    _say("case restricted cylinder")
    r = R_UM / SPACE
    # the prism runs past both z faces so that segments crossing them still
    # see the wall; spins wrap in z
    V, T, I = _prism(r, -1.3, 1.3, NGON)
    sub = D.Substrate(V, T, I, grid="aabb", dims=(26, 26, 26), lo=(-1.3,) * 3, cell=(0.1,) * 3,
                      eps=1e-5)
    n = 200000
    sd = np.float32(np.sqrt(2 * D_ADC * DT) / SPACE)
    sim = D.Simulation(sub, n, dt=DT, std_dev=(sd,) * 4, t2=(np.inf,) * 4, permeability=0.0,
                       mode=D.MODE_PHYSICAL, boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX,
                       seed=21)
    rng = np.random.default_rng(22)
    rin = r * np.cos(np.pi / NGON) * 0.999
    rr = rin * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    pos = np.c_[rr * np.cos(th), rr * np.sin(th), rng.uniform(-1, 1, n)].astype(np.float32)
    sim.set_state(pos=pos)
    sim.label(1.0, 1.0, 1.0)
    comp = sim.get_state()["comp"]
    assert (comp == comp[0]).all(), "every spin starts inside the cylinder"
    col = 1 + int(comp[0])    # signal columns: total, then compartments 0-3
    delta, Delta = 10.0, 20.0
    L, B = int(round(delta / DT)), int(round(Delta / DT))
    sim.moment_reset()
    sim.integrate_moment(1.0, L)
    sim.integrate_moment(0.0, B - L)
    sim.integrate_moment(-1.0, L)
    # mT/m across the axis, where the Gaussian phase approximation holds
    # (S >= 0.93; at 160 and 240 mT/m the walk attenuates 2 % and 4 % more in
    # ln S than the approximation, the sign of the sub-Gaussian phase spread of
    # bounded displacements)
    perp = np.array([40.0, 60.0, 80.0])
    par = np.array([20.0, 40.0])            # along the axis
    g = np.zeros((len(perp) + len(par), 3), np.float32)
    g[:len(perp), 0] = perp * 1e-9 * SPACE
    g[len(perp):, 2] = par * 1e-9 * SPACE
    S = sim.signal_moment(g, PC)
    sim.close()
    assert np.allclose(S[:, 0], S[:, col]), "all spins are in one compartment"
    Reff = A.polygon_area_radius(R_UM, NGON)
    want_perp = np.exp([A.cylinder_gpa_log_signal(G, delta, Delta, D_ADC, Reff) for G in perp])
    want_par = np.exp(A.free_log_signal(par, delta, Delta, D_ADC))
    want = np.r_[want_perp, want_par]
    sigma = np.sqrt((1 - want ** 2) / (2 * n))
    # model allowance across the axis: 1.5 % of |ln S| for the Gaussian phase
    # approximation and the finite step (0.14 um per axis against R = 4 um)
    allow = np.r_[0.015 * np.abs(np.log(want_perp)) * want_perp, np.zeros(len(par))]
    _say(f"  signal {np.round(S[:, col], 5)} want {np.round(want, 5)} "
         f"3 sigma + allowance {np.round(3 * sigma + allow, 5)}")
    assert np.all(np.abs(S[:, col] - want) < 3 * sigma + allow), (S[:, col], want)

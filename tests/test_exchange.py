"""Permeability: the per-hit transmission probability p = kappa sqrt(pi dt / D)
(dsim_amd.substrate.kappa_to_probability, the walk's rule at
spinKernelCpu.cpp:240) against two-compartment exchange.  Spins start inside
a permeable cylinder (a 64-gon prism through a periodic voxel); in the
barrier-limited regime (kappa R / D << 1) the intra-axonal fraction relaxes
as f(t) = f_inf + (1 - f_inf) exp(-k t) with k = kappa S (1/V_in + 1/V_out)
(S, V per unit length) and f_inf = V_in / V.  The exchange rate fitted from
the labels of the walk must match k: on the CPU through the oracle and on
the GPU through dsim_integrate_moment."""
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import substrate as S

DT, DIFF, KAPPA = 0.005, 2.0, 0.05        # ms, um^2/ms, um/ms
R_UM, SPACE, NGON = 2.0, 4.0, 64          # cylinder radius, half voxel (um), polygon
CHUNK = 1000                               # steps between label reads (5 ms)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


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


def _model():
    area = 0.5 * NGON * R_UM ** 2 * np.sin(2 * np.pi / NGON)
    perim = 2 * NGON * R_UM * np.sin(np.pi / NGON)
    vox = (2 * SPACE) ** 2
    k = KAPPA * perim * (1 / area + 1 / (vox - area))
    return k, area / vox


def _start(n, seed):
    rng = np.random.default_rng(seed)
    r = R_UM / SPACE
    rin = r * np.cos(np.pi / NGON) * 0.999
    rr = rin * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    return np.c_[rr * np.cos(th), rr * np.sin(th), rng.uniform(-1, 1, n)].astype(np.float32)


def _fit_rate(ts, fs, finf):
    y = np.log((np.asarray(fs) - finf) / (1 - finf))
    t = np.asarray(ts)
    return float(-(t @ y) / (t @ t))


def test_exchange_rate_oracle():
    k, finf = _model()
    p = S.kappa_to_probability(KAPPA, DIFF, DT)
    assert KAPPA * R_UM / DIFF < 0.1      # barrier-limited
    V, T, I = _prism(R_UM / SPACE, -1.3, 1.3, NGON)
    sd = np.float32(np.sqrt(2 * DIFF * DT) / SPACE)
    w = O.World(V, T, I, grid="aabb", dims=(13, 13, 13), lo=(-1.3,) * 3, cell=(0.2,) * 3,
                eps=1e-5, std_dev=(sd,) * 4, permeability=p, boundary=1, test=1, traversal=1,
                transitions=1)
    n = 20000
    pos = _start(n, 3)
    st = O.State(pos, comp=np.ones(n, np.uint8), fiber=np.zeros(n, np.uint16), phase=None)
    ts, fs = [], []
    for c in range(5):
        O.integrate(w, st, CHUNK, rng=O.RNG_PHILOX, key=(4, 5), step0=c * CHUNK)
        ts.append((c + 1) * CHUNK * DT)
        fs.append(float(np.mean(st.comp == 1)))
    kf = _fit_rate(ts, fs, finf)
    _say(f"oracle: f {np.round(fs, 4)}; fitted k {kf:.4f} /ms, model {k:.4f} /ms")
    assert abs(kf / k - 1) < 0.08


@pytest.mark.gpu
def test_exchange_rate_gpu():
    from dsim_amd import lib as D
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    k, finf = _model()
    p = S.kappa_to_probability(KAPPA, DIFF, DT)
    V, T, I = _prism(R_UM / SPACE, -1.3, 1.3, NGON)
    sub = D.Substrate(V, T, I, grid="aabb", dims=(26, 26, 26), lo=(-1.3,) * 3, cell=(0.1,) * 3,
                      eps=1e-5)
    n = 200000
    sd = np.float32(np.sqrt(2 * DIFF * DT) / SPACE)
    sim = D.Simulation(sub, n, dt=DT, std_dev=(sd,) * 4, permeability=p, mode=D.MODE_PHYSICAL,
                       boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=31)
    sim.set_state(pos=_start(n, 4), comp=np.ones(n, np.uint8), fiber=np.zeros(n, np.uint16))
    sim.moment_reset()
    ts, fs = [], []
    for c in range(6):
        sim.integrate_moment(0.0, CHUNK)
        ts.append((c + 1) * CHUNK * DT)
        fs.append(float(np.mean(sim.get_state()["comp"] == 1)))
    kf = _fit_rate(ts, fs, finf)
    model = finf + (1 - finf) * np.exp(-k * np.array(ts))
    _say(f"gpu: f {np.round(fs, 4)}; model {np.round(model, 4)}; fitted k {kf:.4f} /ms, "
         f"model {k:.4f} /ms")
    assert abs(kf / k - 1) < 0.05
    sigma = np.sqrt(np.asarray(fs) * (1 - np.asarray(fs)) / n)
    assert np.all(np.abs(np.asarray(fs) - model) < 3 * sigma + 0.02 * (1 - finf))

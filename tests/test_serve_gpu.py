"""The server (include/dsim_serve.h) on the GPU: a request's signals are the
same bits alone, in a mix of requests on different substrates, under a budget
that serialises admission, from the background loop, and through the plain
simulation handle (dsim_integrate_moment)."""
import dataclasses

import numpy as np
import pytest

from dsim_amd import lib as D
from dsim_amd import serve as SV
from dsim_amd import substrate as S

pytestmark = pytest.mark.gpu

GAMMA = 42576.0


@pytest.fixture(scope="module")
def envs():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    cfg_a = S.SMALL
    cfg_b = dataclasses.replace(S.SMALL, name="small-b", n_axons=9, seed=7, mean_radius=3.0)
    out = []
    for cfg in (cfg_a, cfg_b):
        env = S.build(cfg, cell_um=0.5, axons=S.pack(cfg))
        g = env["grid"]
        env["sub"] = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], grid="aabb",
                                 dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"])
        out.append(env)
    return out


def _requests(envs):
    rng = np.random.default_rng(3)
    reqs = []
    for k, (e, n, L, B, R, nl, bnd, perm, t2) in enumerate([
            (0, 3000, 300, 800, 100, 6, D.BOUNDARY_PERIODIC, 0.02, (90.0, 70.0, 70.0, 70.0)),
            (1, 700, 150, 500, 0, 3, D.BOUNDARY_PERIODIC, 0.0, (np.inf,) * 4),
            (0, 5000, 200, 600, 50, 4, D.BOUNDARY_CLAMP, 0.05, (80.0, 60.0, 10.0, 60.0)),
            (1, 257, 120, 330, 10, 2, D.BOUNDARY_PERIODIC, 0.01, (90.0, 70.0, 70.0, 70.0))]):
        env = envs[e]
        cfg = env["cfg"]
        grads = (rng.normal(size=(nl, 3)) * 40.0 * 1e-9 * env["space_scale"]).astype(np.float32)
        reqs.append(dict(env=e, n=n, seed=1000 + k, dt=cfg.dt, std_dev=env["std_dev"],
                         t2=t2, permeability=perm,
                         lobes=SV.pgse_lobes(L, B, R), grads=grads, boundary=bnd,
                         extent=(env["xmax"],) * 3,
                         pc=np.float32(2 * np.pi * GAMMA * cfg.dt)))
    return reqs


def _submit(srv, ids, q):
    return srv.submit(ids[q["env"]], q["n"], seed=q["seed"], dt=q["dt"], std_dev=q["std_dev"],
                      t2=q["t2"], permeability=q["permeability"], lobes=q["lobes"],
                      grads=q["grads"], phase_constant=q["pc"], boundary=q["boundary"],
                      label_extent=q["extent"])


def _server(envs, budget=8 << 30, round_steps=100):
    srv = SV.Server(budget, round_steps=round_steps, max_active_spins=1 << 20)
    ids = [srv.add_substrate(e["sub"]) for e in envs]
    return srv, ids


def _plain(envs, q):
    env = envs[q["env"]]
    sim = D.Simulation(env["sub"], q["n"], dt=q["dt"], std_dev=q["std_dev"], t2=q["t2"],
                       permeability=q["permeability"], mode=D.MODE_PHYSICAL,
                       boundary=q["boundary"], rng=D.RNG_PHILOX, seed=q["seed"])
    sim.init_uniform(1.0)
    sim.label(*q["extent"])
    sim.moment_reset()
    for w, k in q["lobes"]:
        if k:
            sim.integrate_moment(w, k)
    out = sim.signal_moment(q["grads"], q["pc"])
    sim.close()
    return out


def _solo(envs, reqs):
    res = []
    for q in reqs:
        srv, ids = _server(envs)
        t = _submit(srv, ids, q)
        srv.run_until_idle()
        res.append(srv.result(t)[0])
        srv.close()
    return res


def test_served_request_equals_plain_path(envs):
    reqs = _requests(envs)
    for q, got in zip(reqs, _solo(envs, reqs)):
        np.testing.assert_array_equal(got, _plain(envs, q))
        assert 0.0 < got[0, 0] <= 1.0


def test_mix_equals_alone(envs):
    reqs = _requests(envs)
    alone = _solo(envs, reqs)
    srv, ids = _server(envs)
    tickets = [_submit(srv, ids, q) for q in reqs]
    srv.run_until_idle()
    for t, want in zip(tickets, alone):
        got, st = srv.result(t)
        np.testing.assert_array_equal(got, want)
        assert st["spin_steps"] > 0 and st["done_s"] >= st["admit_s"] >= st["submit_s"]
    srv.close()


def test_small_first_rounds(envs):
    # a large request of few steps and a small one of many: every round runs
    # both by default, so the large one ends first; with the small-first
    # policy (requests of at most 5 one-wave blocks, 320 spins, run alone) the
    # small one walks in rounds of its own and ends first.  Results equal the
    # solo runs either way.
    base = _requests(envs)
    large = dict(base[0], lobes=SV.pgse_lobes(20, 60, 0))        # 3000 spins, 80 steps
    small = dict(base[3], lobes=SV.pgse_lobes(200, 600, 100))    # 257 spins, 900 steps
    reqs = [large, small]
    alone = _solo(envs, reqs)
    done = {}
    for policy in (0, 5):
        srv = SV.Server(8 << 30, round_steps=100, max_active_spins=1 << 20,
                        small_first_blocks=policy)
        ids = [srv.add_substrate(e["sub"]) for e in envs]
        tickets = [_submit(srv, ids, q) for q in reqs]
        srv.run_until_idle()
        ends = []
        for t, want in zip(tickets, alone):
            got, st = srv.result(t)
            np.testing.assert_array_equal(got, want)
            ends.append(st["done_s"])
        done[policy] = ends
        srv.close()
    assert done[0][0] < done[0][1]      # by default the short large request ends first
    assert done[5][1] < done[5][0]      # small first: the small request ends first


def test_budget_serialises_admission(envs):
    reqs = _requests(envs)
    alone = _solo(envs, reqs)
    # room for the largest request (about 666 KB of state and readout space)
    # but for no two consecutive ones: one request in flight at a time
    srv, ids = _server(envs, budget=700_000)
    tickets = [_submit(srv, ids, q) for q in reqs]
    srv.run_until_idle()
    stats = []
    for t, want in zip(tickets, alone):
        got, st = srv.result(t)
        np.testing.assert_array_equal(got, want)
        stats.append(st)
    for a, b in zip(stats, stats[1:]):
        assert b["admit_s"] >= a["done_s"]
    srv.close()


def test_background_loop(envs):
    reqs = _requests(envs)
    alone = _solo(envs, reqs)
    srv, ids = _server(envs)
    srv.start()
    tickets = [_submit(srv, ids, q) for q in reqs]
    for t, want in zip(tickets, alone):
        srv.wait(t, 120.0)
        np.testing.assert_array_equal(srv.result(t)[0], want)
    srv.stop()
    srv.close()


def test_many_heterogeneous_substrates_in_one_launch():
    # 24 device-built substrates (20-150 axons, mean radius 0.5-3 um, 32-gon,
    # cells 0.3-0.6 um), one request on each: all are admitted together and
    # advanced by the same batched launches, and every result equals the plain
    # path's bits
    rng = np.random.default_rng(11)
    envs, reqs = [], []
    for k in range(24):
        cfg = dataclasses.replace(S.SCALE, name=f"het-{k}", n_axons=int(rng.integers(20, 150)),
                                  mean_radius=float(rng.uniform(0.5, 3.0)), ngon=32,
                                  fraction=0.5, seed=100 + k)
        env = S.build(cfg, cell_um=float(rng.uniform(0.3, 0.6)), axons=S.pack(cfg))
        g = env["grid"]
        env["sub"] = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], grid="aabb",
                                 dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"],
                                 device=True)
        envs.append(env)
        nl = int(rng.integers(2, 5))
        grads = (rng.normal(size=(nl, 3)) * 40.0 * 1e-9 * env["space_scale"]).astype(np.float32)
        reqs.append(dict(env=k, n=int(rng.integers(300, 3000)), seed=2000 + k, dt=cfg.dt,
                         std_dev=env["std_dev"], t2=env["t2"], permeability=0.01,
                         lobes=SV.pgse_lobes(int(rng.integers(60, 120)), 300, 20), grads=grads,
                         boundary=D.BOUNDARY_PERIODIC, extent=(env["xmax"],) * 3,
                         pc=np.float32(2 * np.pi * GAMMA * cfg.dt)))
    srv, ids = _server(envs)
    tickets = [_submit(srv, ids, q) for q in reqs]
    srv.run_until_idle()
    stats = []
    for t, q in zip(tickets, reqs):
        got, st = srv.result(t)
        np.testing.assert_array_equal(got, _plain(envs, q))
        stats.append(st)
    # every request was admitted before any finished: all shared the launches
    assert max(s["admit_s"] for s in stats) <= min(s["done_s"] for s in stats)
    srv.close()
    for e in envs:
        e["sub"].close()

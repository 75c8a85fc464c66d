"""The walk's work counters (dsim_walk_stats_*): the counting build changes
no result, its step and membrane-hit counts equal the oracle's on injected
steps, and the counts obey the walk's own bookkeeping (every step is either
a clearance shortcut or searched at least once; every triangle test follows
a plane crossing; every hit is a triangle test that passed)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D

pytestmark = pytest.mark.gpu

DT = 0.005


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _setup(ref_inputs, n, boundary, sort_interval=0, seed=1):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 6.0)
    dims, lo, cell = (37, 29, 13), (-1.2, -1.2, -1.2), (2.4 / 37, 2.4 / 29, 2.4 / 13)
    std, t2 = (0.03, 0.03, 0.01, 0.03), (80.0, 70.0, 7.5, 70.0)
    w = O.World(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=dims, lo=lo,
                cell=cell, eps=1e-5, std_dev=std, t2=t2, dt=DT, permeability=0.2)
    w.set(test=1, traversal=1, transitions=1, boundary=boundary)
    sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=dims, lo=lo,
                      cell=cell, eps=1e-5)
    sim = D.Simulation(sub, n, dt=DT, std_dev=std, t2=t2, permeability=0.2,
                       mode=D.MODE_PHYSICAL, boundary=boundary, rng=D.RNG_PHILOX, seed=seed,
                       sort_interval=sort_interval)
    rng = np.random.default_rng(9)
    pos = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    comp, fib, _, _ = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"], plus_x=True)
    return w, sub, sim, pos, comp, fib


def _consistent(st, n_steps):
    assert st["steps"] == n_steps
    assert st["search_lanes"] >= st["steps"] - st["shortcuts"]
    assert st["plane_tests"] >= st["tri_tests"] >= st["hits"]
    assert st["tri_tests"] >= st["fp64_tests"]
    assert st["live_lanes"] >= st["search_lanes"] + st["deferred"]
    assert st["trips"] <= st["live_lanes"]


@pytest.mark.parametrize("boundary", [0, 1])
def test_counts_match_oracle_on_injected_steps(ref_inputs, boundary):
    _say(f"case stats-injected boundary={boundary}")
    n, iters = 3000, 48
    w, sub, sim, pos, comp, fib = _setup(ref_inputs, n, boundary)
    q = np.random.default_rng(10).normal(size=(iters, n, 4)).astype(np.float32)
    q[..., 3] = np.random.default_rng(11).uniform(0, 1, (iters, n)).astype(np.float32)
    st = O.State(pos, comp=comp, fiber=fib, moment=True)
    ost = np.zeros(3, np.uint64)
    O.integrate(w, st, iters, rng=O.RNG_INJECT, inject=q, weight=1.0, stats=ost)
    sim.set_state(pos=pos, comp=comp, fiber=fib)
    sim.moment_reset()
    sim.walk_stats(True)
    sim.integrate_moment_injected(1.0, q)
    g = sim.get_state()
    c = sim.walk_stats()
    sim.walk_stats(False)
    np.testing.assert_array_equal(g["pos"], st.pos)
    np.testing.assert_array_equal(g["comp"], st.comp)
    _say(f"  {c}")
    assert c["steps"] == int(ost[0]) == n * iters
    assert c["hits"] == int(ost[2]) > 300
    assert c["waves"] == (n + 63) // 64
    _consistent(c, n * iters)


def test_counting_build_changes_nothing(ref_inputs):
    _say("case stats-philox")
    n, iters = 20000, 300
    out = []
    for counting in (False, True):
        w, sub, sim, pos, comp, fib = _setup(ref_inputs, n, 1, sort_interval=64, seed=5)
        sim.set_state(pos=pos, comp=comp, fiber=fib)
        sim.moment_reset()
        if counting:
            sim.walk_stats(True)
        sim.integrate_moment(1.0, iters)
        g = sim.get_state()
        mom, _ = sim.get_moment()
        out.append((g, mom, sim.walk_stats()))
    (g0, m0, _), (g1, m1, c) = out
    for k in ("pos", "mag", "comp", "fiber"):
        np.testing.assert_array_equal(g0[k], g1[k])
    np.testing.assert_array_equal(m0, m1)
    _say(f"  {c}")
    _consistent(c, n * iters)
    assert c["waves"] == 5 * ((n + 63) // 64)  # five launches of at most 64 steps


def test_counting_refused_for_mwc(ref_inputs):
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 6.0)
    sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"])
    sim = D.Simulation(sub, 64, dt=DT, std_dev=(0.03,) * 4, mode=D.MODE_PHYSICAL,
                       rng=D.RNG_MWC)
    sim.walk_stats(True)
    with pytest.raises(D.DsimError):
        sim.integrate(np.zeros(3, np.float32), 0.0, 1)


def test_persistent_walk_equals_fixed_blocks(ref_inputs):
    # 6e5 spins: at least two per resident one-wave workgroup, and launches of
    # 1024 steps (the persistent walk takes launches of 768 or more), so the
    # persistent walk (lanes taking their next spin from a per-XCD queue) runs,
    # twice, with a reorder between; its results and counts equal the
    # fixed-block walk's
    _say("case persistent walk")
    n, iters = 600_000, 2048
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 6.0)
    dims, lo, cell = (37, 29, 13), (-1.2, -1.2, -1.2), (2.4 / 37, 2.4 / 29, 2.4 / 13)
    sub = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], grid="aabb", dims=dims, lo=lo,
                      cell=cell, eps=1e-5)
    out = []
    for fixed in (True, False):
        sim = D.Simulation(sub, n, dt=DT, std_dev=(0.03, 0.03, 0.01, 0.03),
                           t2=(80.0, 70.0, 7.5, 70.0), permeability=0.2, mode=D.MODE_PHYSICAL,
                           boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=7,
                           sort_interval=1024, fixed_blocks=fixed)
        sim.init_uniform(0.9)
        sim.label(m["xmax"], m["ymax"], m["zmax"])
        sim.moment_reset()
        sim.walk_stats(True)
        sim.integrate_moment(1.0, iters)
        g = sim.get_state()
        mom, _ = sim.get_moment()
        out.append((g, mom, sim.walk_stats()))
        sim.close()
    (g0, m0, c0), (g1, m1, c1) = out
    for k in ("pos", "mag", "comp", "fiber"):
        np.testing.assert_array_equal(g0[k], g1[k])
    np.testing.assert_array_equal(m0, m1)
    _say(f"  fixed {c0}\n  persistent {c1}")
    assert c0["steps"] == c1["steps"] == n * iters
    assert c0["hits"] == c1["hits"] > 0
    assert c1["waves"] < c0["waves"]
    _consistent(c1, n * iters)

"""Device-built substrates (dsim_substrate_create_device) against the host
build (dsim_substrate_create): triangle helpers, grid lists, entry counts and
clearance must be identical, and a walk on either must give identical
states."""
import os
import sys
import time

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D
from dsim_amd import substrate as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _both(m, **grid):
    a = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], **grid)
    b = D.Substrate(m["vertices"], m["trgls"], m["tri_info"], device=True, **grid)
    return a, b


def _same(a, b, clearance):
    sa, ta, ma = a.grid()
    sb, tb, mb = b.grid()
    np.testing.assert_array_equal(sb, sa)
    np.testing.assert_array_equal(tb, ta)
    assert mb == ma
    np.testing.assert_array_equal(b.helpers(), a.helpers())
    if clearance:
        np.testing.assert_array_equal(b.clearance(), a.clearance())
    return len(ta)


CASES = [("sim.fibers_1_r004_f09", 6.0), ("sim.fibers_1_r0064_f05", 8.0), ("sim.triangles_curved", 1.0)]


@pytest.mark.parametrize("name,scale", CASES)
def test_reference_grid_equals_host_build(ref_inputs, name, scale):
    _say(f"case device grid reference {name}")
    m = O.parse_mesh(os.path.join(ref_inputs, name), scale)
    a, b = _both(m)
    assert _same(a, b, clearance=False) > 0


@pytest.mark.parametrize("name,scale", CASES)
def test_aabb_grid_equals_host_build(ref_inputs, name, scale):
    _say(f"case device grid aabb {name}")
    m = O.parse_mesh(os.path.join(ref_inputs, name), scale)
    dims, lo, cell = (37, 29, 13), (-1.2, -1.2, -1.2), (2.4 / 37, 2.4 / 29, 2.4 / 13)
    a, b = _both(m, grid="aabb", dims=dims, lo=lo, cell=cell, eps=1e-5)
    assert _same(a, b, clearance=True) > 0


def test_generated_substrate_and_walk_equal_host_build():
    # 300 packed axons, 64-gon prisms, 0.2 um cells: the bench's substrate
    # family at a size the host build finishes quickly
    _say("case device grid generated substrate")
    cfg = S.SCALE
    rng = np.random.default_rng(5)
    ax = S.make_axons(300, cfg.gamma_k, cfg.mean_radius, rng, fraction=0.6, ngon=64)
    env = S.build(cfg, cell_um=0.2, axons=ax)
    g = env["grid"]
    kw = dict(grid="aabb", dims=g["dims"], lo=g["lo"], cell=g["cell"], eps=g["eps"])
    t0 = time.perf_counter()
    a = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], **kw)
    t1 = time.perf_counter()
    b = D.Substrate(env["vertices"], env["triangles"], env["tri_info"], device=True, **kw)
    t2 = time.perf_counter()
    _say(f"  {len(env['triangles'])} triangles, grid {g['dims']}: host build {t1 - t0:.3f} s, "
         f"device build {t2 - t1:.3f} s")
    assert _same(a, b, clearance=True) > 0
    states = []
    for sub in (a, b):
        sim = D.Simulation(sub, 20000, dt=cfg.dt, std_dev=env["std_dev"], t2=env["t2"],
                           permeability=env["permeability"], mode=D.MODE_PHYSICAL,
                           boundary=D.BOUNDARY_PERIODIC, rng=D.RNG_PHILOX, seed=11)
        sim.init_uniform(1.0)
        sim.label(env["xmax"], env["xmax"], env["xmax"])
        sim.moment_reset()
        sim.integrate_moment(1.0, 200)
        states.append(sim.get_state())
        sim.close()
    for k in ("pos", "mag", "comp", "fiber"):
        np.testing.assert_array_equal(states[1][k], states[0][k])


def test_empty_mesh_and_partial_grid_equal_host_build(ref_inputs):
    _say("case device grid edge cases")
    # no triangles: empty lists, clearance 1e30 everywhere
    m = {"vertices": np.zeros((3, 3), np.float32), "trgls": np.zeros((0, 3), np.uint32),
         "tri_info": np.zeros((0, 3), np.uint32)}
    a, b = _both(m, grid="aabb", dims=(5, 4, 3), lo=(-1.0,) * 3, cell=(0.4, 0.5, 0.6), eps=0.0)
    assert _same(a, b, clearance=True) == 0
    assert (b.clearance() == np.float32(1e30)).all()
    # a grid over part of a mesh: triangles outside it are dropped, those
    # straddling its faces are kept
    m = O.parse_mesh(os.path.join(ref_inputs, "sim.fibers_1_r004_f09"), 6.0)
    a, b = _both(m, grid="aabb", dims=(17, 11, 7), lo=(-0.3, -0.2, -0.5), cell=(0.03, 0.04, 0.05),
                 eps=2e-5)
    assert _same(a, b, clearance=True) > 0

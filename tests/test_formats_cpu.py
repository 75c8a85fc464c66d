"""Product input/output formats (include/dsim_io.h) against the oracle's
restatement of the reference readers, the reference's own sample files and
closed forms."""
import os
import re

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D

SAMPLES = [("sim.fibers_1_r004_f09", 60.0), ("sim.fibers_1_r0064_f05", 4.0),
           ("sim.triangles_curved", 1.0)]


@pytest.mark.parametrize("name,scale", SAMPLES)
def test_mesh_reader_matches_reference_reader(ref_inputs, name, scale):
    path = os.path.join(ref_inputs, name)
    p = D.read_mesh(path, scale)
    o = O.parse_mesh(path, scale)
    assert o["rc"] == 0
    for k in ("n_axon", "n_myelin", "n_glia", "n_fibers"):
        assert p[k] == o[k], k
    assert p["max_triangles_on_surface"] == o["max_tri_on_surface"]
    np.testing.assert_array_equal(p["vertices"], o["vertices"])
    np.testing.assert_array_equal(p["triangles"], o["trgls"])
    np.testing.assert_array_equal(p["tri_info"], o["tri_info"])
    assert (p["xmax"], p["ymax"], p["zmax"]) == (o["xmax"], o["ymax"], o["zmax"])


def test_mesh_header_is_tolerated(ref_inputs):
    # the reference counts the header word Max_triangles_on_membrane as a
    # myelin section and rejects the file; the product skips the header
    m = D.read_mesh(os.path.join(ref_inputs, "sim.triangles_8f_1g"), 1.0)
    assert (m["n_vertices"], m["n_triangles"]) == (2736, 5400)
    assert (m["n_axon"], m["n_myelin"], m["n_glia"], m["n_fibers"]) == (8, 8, 1, 9)
    assert m["max_triangles_on_surface"] == 300
    assert set(np.unique(m["tri_info"][:, 1])) == {0, 1, 2}


def test_mesh_errors(tmp_path, ref_inputs):
    with pytest.raises(D.DsimError) as e:
        D.read_mesh(tmp_path / "missing", 1.0)
    assert e.value.code == -6
    bad = tmp_path / "unpaired"
    bad.write_text("V\n0 0 0\n1 0 0\n0 1 0\nT\n0 1 2\nA0\n0\n")
    with pytest.raises(D.DsimError) as e:
        D.read_mesh(bad, 1.0)
    assert e.value.code == -7
    oob = tmp_path / "oob"
    oob.write_text("V\n0 0 0\n1 0 0\n0 1 0\nT\n0 1 3\n")
    with pytest.raises(D.DsimError) as e:
        D.read_mesh(oob, 1.0)
    assert e.value.code == -7


def test_mesh_write_round_trip(tmp_path, ref_inputs):
    m = D.read_mesh(os.path.join(ref_inputs, "sim.triangles_8f_1g"), 1.0)
    out = tmp_path / "rt.fibers"
    D.write_mesh(out, m["vertices"], m["triangles"], m["tri_info"], 1.0)
    r = D.read_mesh(out, 1.0)
    np.testing.assert_array_equal(r["vertices"], m["vertices"])
    np.testing.assert_array_equal(r["triangles"], m["triangles"])
    np.testing.assert_array_equal(r["tri_info"], m["tri_info"])
    # and the reference reader accepts what the writer produces (no header)
    o = O.parse_mesh(str(out), 1.0)
    assert o["rc"] == 0 and o["n_tri"] == 5400


def test_grads_reader_matches_reference_reader(ref_inputs):
    path = os.path.join(ref_inputs, "sim.grads")
    np.testing.assert_array_equal(D.read_grads(path), O.parse_grads(path))


def test_config_reader(ref_inputs):
    c = D.read_config(os.path.join(ref_inputs, "sim.cfg"))
    # dSim3/sim.cfg:1-34
    assert (c.stepsPerUpdate, c.useGpu, c.useDisplay, c.writeOut) == (10, 0, 0, 1)
    assert c.numSpins == 30000
    assert (c.gyroMagneticRatio, c.timeStep) == (42576.0, np.float32(0.005))
    assert (c.extraAdc, c.intraAdc, c.myelinAdc) == (np.float32(2.1), np.float32(2.1), np.float32(0.1))
    assert (c.extraT2, c.intraT2, c.myelinT2) == (80.0, 80.0, 7.5)
    assert c.permeability == np.float32(0.000001)
    assert (c.spaceScale, c.startBox) == (60.0, np.float32(0.6))
    assert c.gradsFile == b"sim.grads" and c.outFile == b"results_spl_g65.m"
    assert c.fiberFile == b"fibers/splenium_085_g65.fmf"
    # command-line overrides (diffusion.cpp:219-268)
    c = D.read_config(os.path.join(ref_inputs, "sim.cfg"),
                      ["-numSpins=500", "-perm=0.25", "-gpu", "-nw", "-fiberFile=x.fibers",
                       "--startBox=0.3", "-stepsPerUpdate=7"])
    assert (c.numSpins, c.useGpu, c.writeOut, c.stepsPerUpdate) == (500, 1, 0, 7)
    assert c.permeability == np.float32(0.25) and c.startBox == np.float32(0.3)
    assert c.fiberFile == b"x.fibers"
    d = D.read_config()
    assert (d.timeStep, d.numSpins, d.spaceScale) == (np.float32(0.01), 30000, 100.0)


def test_config_syntax(tmp_path):
    p = tmp_path / "a.cfg"
    p.write_text('/* c */ sim: { timestep = 0.002; # dSim2 name\n gyromagneticRatio = 1e3;\n'
                 ' outFile = "a" "b.m"; };\nfibers = { file = "f.fibers"; };\napp: {useGpu = TRUE;};')
    c = D.read_config(p)
    assert c.timeStep == np.float32(0.002) and c.gyroMagneticRatio == 1000.0
    assert c.outFile == b"ab.m" and c.fiberFile == b"f.fibers" and c.useGpu == 1
    q = tmp_path / "bad.cfg"
    q.write_text("sim: { numSpins = 5; ")
    with pytest.raises(D.DsimError) as e:
        D.read_config(q)
    assert e.value.code == -7


@pytest.mark.parametrize("args", [(10.0, 12.0669, 1.0, 0.005, 10), (10.0, 40.0, 10.0, 0.005, 1),
                                  (10.0, 5.0, 1.0, 0.5, 1), (10.0, 40.0, 0.0, 0.5, 1),
                                  (3.3, 7.1, 2.2, 0.01, 3)])
def test_pgse_schedule_matches_reference(args):
    s, lbr = D.pgse_schedule(*args)
    so, lbro = O.pgse_schedule(*args)
    assert lbr == lbro
    np.testing.assert_array_equal(s, so)


def _signals_for_tensor(Dt, rows, s0=1.0):
    out = []
    for r in rows:
        g = np.asarray(r[3:6], np.float64)
        nrm = np.linalg.norm(g)
        if nrm == 0:
            out.append(s0)
            continue
        gh = g / nrm
        b = (2 * np.pi * 42576.0 * nrm * 1e-9 * r[0]) ** 2 * (r[1] - r[0] / 3.0)
        out.append(s0 * np.exp(-b * gh @ Dt @ gh))
    return np.array(out, np.float32)


def test_tensor_fit_recovers_closed_form(ref_inputs):
    rows = O.parse_grads(os.path.join(ref_inputs, "sim.grads"))
    R = np.array([[0.9, 0.3, 0.1], [-0.2, 0.9, 0.3], [0.1, -0.3, 0.95]])
    Q, _ = np.linalg.qr(R)
    evals = np.array([0.3, 0.5, 1.7])
    Dt = Q @ np.diag(evals) @ Q.T
    fit = D.tensor_fit(rows, _signals_for_tensor(Dt, rows))
    np.testing.assert_allclose(fit["tensor"], Dt, atol=2e-4)
    np.testing.assert_allclose(fit["eigval"], evals, rtol=1e-3)
    assert fit["md"] == pytest.approx(evals.mean(), rel=1e-3)
    fa = np.sqrt(1.5) * np.sqrt(((evals - evals.mean()) ** 2).sum()) / np.sqrt((evals ** 2).sum())
    assert fit["fa"] == pytest.approx(fa, rel=2e-3)
    assert fit["rd"] == pytest.approx((0.3 + 0.5) / 2, rel=1e-3)
    v = fit["eigvec"][:, 2]
    assert abs(abs(v @ Q[:, 2]) - 1) < 1e-3


def test_write_m_layout(tmp_path, ref_inputs):
    rows = O.parse_grads(os.path.join(ref_inputs, "sim.grads"))
    c = D.read_config(os.path.join(ref_inputs, "sim.cfg"))
    sig = np.tile(np.linspace(1, 0.2, 5, dtype=np.float32), (len(rows), 1))
    fit = D.tensor_fit(rows, _signals_for_tensor(np.diag([1.0, 1.0, 2.0]), rows))
    out = tmp_path / "r.m"
    D.write_m(out, c, rows, sig, [fit] * 5)
    D.write_m(out, c, rows, sig, [fit] * 5)  # appended, like fopen("at")
    text = out.read_text()
    assert text.count("if (exist('m','var')), m=m+1; else, m=1; end") == 2
    assert "numSpins(m) = 30000;" in text and "spaceScale(m) = 60;" in text
    assert "fibersFile{m} = 'fibers/splenium_085_g65.fmf';" in text
    assert "%%%%%%%%%%%%%%%%% \n% Input parameters \n%%%%%%%%%%%%%%%%%\n" in text
    assert len(re.findall(r"^tensor\{m\}\{[1-4]\} = \[", text, re.M)) == 8
    assert "tensor{m}{5}" not in text  # the reference writes indices 1..4 only
    assert text.count("s=s+1;") == 26
    assert "dwGrads{m}(:,s) = [80,80,0];" in text
    assert "mrSigCompartments{m}(4,s) = 0.2;" in text
    assert "Delta{m}(s) = 12.0669;" in text

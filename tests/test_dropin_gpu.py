"""The drop-in end to end on the GPU: the `dsim` command line (the twin of the
reference's `diffusion` program, diffusion.cpp:1068-1184) on a sample mesh
and the reference's sample sim.grads, its .m output against the layout of
the reference's own sample output (matlab_tests/data_outputs/testfile.m,
written from the same sim.grads) and its signals against the oracle; and the
SpinSystem call sequence of the reference's host (dsim_example: initFibers,
populateCubes, build, setArray, findSpinsInFibers, getArray, updateSpins,
getMrSignal) against the oracle."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from dsim_amd import lib as D

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "golden", "reference_inputs")
SAMPLE_OUT = os.path.join(ROOT, "tests", "golden", "reference_outputs", "testfile.m")
BIN = os.path.join(ROOT, "dsim_amd", "bin")
MESH = os.path.join(INPUTS, "sim.fibers_1_r004_f09")
SPACE = 6.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if D.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")


def _cfg(tmp, n, out, extra=""):
    text = f"""app: {{
    stepsPerUpdate = 10;
    useGpu = true;
    useDisplay = false;
    writeOut = true;
}};
sim: {{
    numSpins = {n};
    gyroMagneticRatio = 42576.0;
    timeStep = 0.005;
    extraAdc = 2.1;
    intraAdc = 2.1;
    myelinAdc = 0.1;
    extraT2 = 80.0;
    intraT2 = 80.0;
    myelinT2 = 7.5;
    permeability = 0.000001;
    spaceScale = {SPACE};
    startBox = 0.6;
    gradsFile = "{os.path.join(INPUTS, 'sim.grads')}";
    outFile = "{out}";
{extra}}};
fibers: {{
    fiberFile = "{MESH}";
}};
"""
    p = os.path.join(tmp, "sim.cfg")
    with open(p, "w") as f:
        f.write(text)
    return p


def _keys(lines):
    """(key, index text, value text) of each assignment line `key...(...) = v;`."""
    out = []
    for ln in lines:
        m = re.match(r"^\s*([A-Za-z]\w*)([({][^=]*)?=\s*(.*?);?\s*$", ln)
        if m and not ln.lstrip().startswith("%"):
            out.append((m.group(1), (m.group(2) or "").strip(), m.group(3).strip()))
    return out


def _blocks(text):
    """Per gradient run: the assignments after each `s=s+1;`."""
    parts = text.split("s=s+1;")
    return [_keys(p.splitlines()) for p in parts[1:]]


def _g(x):
    return "%g" % x


def _signals(text, key):
    return [float(v) for k, i, v in _keys(text.splitlines()) if k == key]


def _oracle_signals(n, seed):
    """The reference-mode oracle on the same mesh, physics and sequence, from
    an independent uniform draw in the start box (diffusion.cpp:564-608)."""
    m = O.parse_mesh(MESH, SPACE)
    s2 = np.float32(SPACE) * np.float32(SPACE)
    dt = np.float32(0.005)
    adc = np.array([2.1, 2.1, 0.1, 2.1], np.float32)  # extra, intra, myelin, glia interior
    std = np.sqrt(np.float32(2) * adc / s2 * dt).astype(np.float32)
    w = O.World(m["vertices"], m["trgls"], m["tri_info"], std_dev=std,
                t2=(80.0, 80.0, 7.5, 80.0), dt=0.005, permeability=1e-6)
    grads = O.parse_grads(os.path.join(INPUTS, "sim.grads"))
    rng = np.random.default_rng(seed)
    pc = np.float32(2 * np.pi * 42576.0 * 0.005)
    out = []
    for row in grads:
        pos = rng.uniform(-0.6, 0.6, (n, 3)).astype(np.float32)
        comp, fib, _, _ = w.find_spins(pos, m["xmax"], m["ymax"], m["zmax"])
        seeds = rng.integers(1, 2 ** 32, (n, 2), dtype=np.uint64).astype(np.uint32)
        st = O.State(pos, comp=comp, fiber=fib, seed=seeds)
        sign, (L, B, R, ro) = O.pgse_schedule(row[0], row[1], row[2], 0.005, 10)
        g = (row[3:6] * np.float32(1e-9) * np.float32(SPACE)).astype(np.float32)
        k = 0
        while k < len(sign):
            k1 = k
            while k1 < len(sign) and sign[k1] == sign[k]:
                k1 += 1
            O.integrate(w, st, (k1 - k) * 10, rng=O.RNG_MWC, grad=(g * sign[k]).astype(np.float32),
                        phase_constant=pc)
            k = k1
        out.append(O.mr_signal(st.mag, st.phase, st.comp))
    return np.array(out)


@pytest.mark.parametrize("shared", [False, True])
def test_cli_output_layout_and_signals(tmp_path, shared):
    n = 20000
    out = str(tmp_path / "out.m")
    cfg = _cfg(str(tmp_path), n, out)
    _say(f"dsim CLI: {n} spins, 13 lines, shared walk {shared}")
    cmd = [os.path.join(BIN, "dsim"), f"-config={cfg}", "-seed=7"] + (["-sharedWalk"] if shared else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = open(out).read()
    ref = open(SAMPLE_OUT).read()
    gl = [ln for ln in got.splitlines() if ln.strip()]
    rl = [ln for ln in ref.splitlines() if ln.strip()]
    # header (diffusion.cpp:668-670)
    assert gl[0].startswith("% * * * ") and rl[0].startswith("% * * * ")
    assert gl[1] == rl[1] == "if (exist('m','var')), m=m+1; else, m=1; end"
    assert gl[2] == rl[2] == "s=0;"
    # input parameters: the sample's keys in the sample's order
    ref_in = [k for k, _, _ in _keys(ref.split("% Simulation data")[0].splitlines())]
    got_in = [k for k, _, _ in _keys(got.split("% Simulation data")[0].splitlines())]
    assert got_in == ref_in
    vals = {k: v for k, _, v in _keys(got.split("% Simulation data")[0].splitlines())}
    assert vals["numSpins"] == str(n) and vals["timeStep"] == "0.005"
    assert vals["spaceScale"] == _g(SPACE) and vals["permeability"] == "1e-06"
    # per gradient run: s=s+1 then delta, Delta, readOut, dwGrads as the sample
    # (same sim.grads, so the same %g text), then the signals
    gb, rb = _blocks(got), _blocks(ref)
    assert len(gb) == len(rb) == 13
    for b_got, b_ref in zip(gb, rb):
        assert [k for k, _, _ in b_got[:4]] == ["delta", "Delta", "readOut", "dwGrads"]
        assert b_got[:4] == b_ref[:4]
        assert [k for k, _, _ in b_got[4:]] == ["mrSigTotal"] + ["mrSigCompartments"] * 4
        for _, _, v in b_got[4:]:
            assert v == _g(float(v))  # %g text
    # the tensor fits of diffusion.cpp:690-698: getNumCompartments() = 4
    # entries (the whole volume, then compartments 0-2)
    assert [i for k, i, _ in _keys(got.splitlines()) if k == "tensor"] == \
           ["{m}{%d}" % c for c in range(1, 5)]
    # signals against the oracle (an independent draw of the same size)
    sig = np.array(_signals(got, "mrSigTotal"))
    ref_sig = _oracle_signals(n, 11)[:, 0]
    sigma = np.sqrt(np.maximum(1 - ref_sig ** 2, 0.01) / (2 * n)) * np.sqrt(2)
    _say("  cli " + " ".join(f"{v:.4f}" for v in sig))
    _say("  orc " + " ".join(f"{v:.4f}" for v in ref_sig))
    assert np.all(np.abs(sig - ref_sig) < 4 * sigma + 2e-3), (sig, ref_sig)
    # the reference's own sample (spaceScale 20, its own mesh) is not this
    # run; only its layout is compared


def test_spin_system_sequence_matches_oracle(tmp_path):
    # the reference host's call sequence through dsim::SpinSystem
    n, iters = 4000, 200
    out = str(tmp_path / "state.bin")
    r = subprocess.run([os.path.join(BIN, "dsim_example"), MESH, str(SPACE), str(n), str(iters),
                        out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(out, "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == n
    off = 4

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off).copy()
        off += a.nbytes
        return a

    def state():
        pos = take(np.float32, 3 * n).reshape(n, 3)
        sp = take(D.SPIN_DTYPE, n)
        seeds = take(np.uint32, 2 * n).reshape(n, 2)
        return pos, sp, seeds

    pos0, sp0, seeds0 = state()
    pos1, sp1, seeds1 = state()
    sig = take(np.float64, 5)
    m = O.parse_mesh(MESH, SPACE)
    s2 = np.float32(SPACE) * np.float32(SPACE)
    dt = np.float32(0.005)
    adc = np.array([2.1, 2.1, 0.1, 2.1], np.float32)
    std = np.sqrt(np.float32(2) * adc / s2 * dt).astype(np.float32)
    w = O.World(m["vertices"], m["trgls"], m["tri_info"], std_dev=std,
                t2=(80.0, 80.0, 7.5, 80.0), dt=0.005, permeability=0.2)
    # findSpinsInFibers: labels equal the oracle's
    comp, fib, _, _ = w.find_spins(pos0, m["xmax"], m["ymax"], m["zmax"])
    np.testing.assert_array_equal(sp0["compartmentType"], comp)
    np.testing.assert_array_equal(sp0["insideFiber"], fib)
    assert len(np.unique(comp)) >= 2
    # updateSpins under (80, 80, 0) mT/m with the MWC seeds
    st = O.State(pos0, comp=comp, fiber=fib, seed=seeds0)
    g = (np.array([80.0, 80.0, 0.0], np.float32) * np.float32(1e-9) * np.float32(SPACE)).astype(np.float32)
    pc = np.float32(2 * np.pi * 42576.0 * 0.005)
    O.integrate(w, st, iters, rng=O.RNG_MWC, grad=g, phase_constant=pc)
    np.testing.assert_array_equal(seeds1, st.seed)  # integer MWC stream exact
    close = np.all(np.abs(pos1 - st.pos) < 1e-5, axis=1)
    assert close.mean() > 0.97  # the rest part at a wall after last-bit Box-Muller differences
    assert np.mean(sp1["compartmentType"] == st.comp) > 0.97
    ref = O.mr_signal(st.mag, st.phase, st.comp)
    assert abs(sig[0] - ref[0]) < 5e-3, (sig, ref)

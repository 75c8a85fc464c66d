"""Builds libdsim.so (HIP kernels for gfx950 + C++ host runtime) in-tree.

    python -m dsim_amd.build

Objects go to dsim_amd/_build/, the library to dsim_amd/libdsim.so and the
command-line driver to dsim_amd/bin/dsim.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdsim.so")
BIN = os.path.join(HERE, "bin", "dsim")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

# Device code: contraction off (every fused multiply-add is an explicit fmaf)
# and packed-FP32 instructions disabled.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-bitwise-instead-of-logical", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
             "-I" + os.path.join(ROCM, "include"), "-I" + os.path.join(ROOT, "include"),
             "-Wall", "-Wno-misleading-indentation", "-Wno-unused-result", "-pthread"]

HIP_SOURCES = ["kernels.hip", "grid.hip", "nalloc.hip"]
CXX_SOURCES = ["host.cpp", "sim.cpp", "formats.cpp", "serve.cpp", "phase.cpp"]
MAIN_SOURCES = ["spin_system.cpp", "dsim_main.cpp"]
EXAMPLE_SOURCES = ["spin_system.cpp", "example_spin_system.cpp"]
EXAMPLE = os.path.join(HERE, "bin", "dsim_example")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    main_objs = []
    for s in MAIN_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        main_objs.append(o)
    ex_objs = [os.path.join(BUILD, s + ".o") for s in EXAMPLE_SOURCES]
    jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, EXAMPLE_SOURCES[1]), "-o",
                 ex_objs[1]])
    with ThreadPoolExecutor(max_workers=4) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs, "-pthread"])
    os.replace(tmp, LIB)
    if not main_objs:
        return LIB
    link = ["-L" + HERE, "-ldsim", "-Wl,-rpath,$ORIGIN/..", "-L" + os.path.join(ROCM, "lib"),
            "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-pthread"]
    _run(["g++", "-o", BIN, *main_objs, *link])
    _run(["g++", "-o", EXAMPLE, *ex_objs, *link])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host code (the
host half of libdsim and the oracle) over the reference's sample inputs:
tests/sanitize/run.sh builds both drivers with -fsanitize=address,undefined
(recovery off, leak checking on) and a report fails the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _have_asan():
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    return r.returncode == 0 and os.path.isabs(r.stdout.strip())


@pytest.mark.skipif(shutil.which("g++") is None or not _have_asan(), reason="no g++ with ASan")
def test_host_code_under_asan_ubsan(tmp_path):
    r = subprocess.run(["sh", os.path.join(HERE, "sanitize", "run.sh"), str(tmp_path)],
                       capture_output=True, text=True, timeout=600)
    log = r.stdout + r.stderr
    assert r.returncode == 0, log[-4000:]
    assert "host driver: ok" in log and "oracle driver: ok" in log
    assert "ERROR: AddressSanitizer" not in log and "runtime error" not in log

"""CPU check of the saved-boundary ring of the RTM source wavefield
(csrc/mnad_ring.h): tests/host/ring_host.cpp is built with g++ (plain and
under the undefined-behaviour sanitizer) and run on the host.  Each case
prints "ok <name>" or "FAIL <name>: <reason>"; what it checks is listed at the
top of that file (ring layout one to one, the 16-bit codes against
independent definitions, bit-exact reverse decoding of the closed-loop
predictive code, its error bound, and that a higher predictor order helps)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.parametrize("san", ["plain", "ubsan"])
def test_ring_layout_and_code(tmp_path, san):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "ring_host")
    cmd = [gxx, "-O2", "-std=c++17", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "mnad-modeling-rtm_amd", "csrc"),
           os.path.join(ROOT, "tests", "host", "ring_host.cpp"), "-o", exe]
    if san == "ubsan":
        cmd[1:1] = ["-fsanitize=undefined", "-fno-sanitize-recover=undefined"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    lines = [ln for ln in r.stdout.strip().splitlines() if not ln.startswith("#")]
    assert len(lines) == 13, r.stdout
    bad = [ln for ln in lines if not ln.startswith("ok ")]
    assert not bad and r.returncode == 0, "\n".join(bad)

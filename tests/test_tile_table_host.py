"""CPU check of the launch tile table that the runtime builds for every step
launch (build_tiles in csrc/mnad_runtime.hip): tests/host/tiles_host.cpp is
compiled against the runtime source alone, with the kernel launchers stubbed,
and run on the host. Each case prints "ok <name>" or "FAIL <name>: <reason>";
what it checks is listed at the top of that file (the tile tables of the
one-step kernels, the block heights of the resident kernel, and the host pass
over a velocity model that validates it and keys the model cache)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


# The host code also runs under the undefined-behaviour sanitizer (host pass
# only; no GPU code is built with it): signed overflow in the model scan,
# shifts and out-of-range conversions abort the run.
UBSAN = ["-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=undefined"]


@pytest.mark.parametrize("san", ["plain", "ubsan"])
def test_tile_table_covers_balances_and_orders(tmp_path, san):
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "tiles_host")
    csrc = os.path.join(ROOT, "mnad-modeling-rtm_amd", "csrc")
    cmd = [HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           *(UBSAN if san == "ubsan" else []),
           "-I", os.path.join(ROOT, "include"), "-I", csrc,
           os.path.join(ROOT, "tests", "host", "tiles_host.cpp"),
           os.path.join(csrc, "mnad_runtime.hip"), "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 23, r.stdout
    bad = [ln for ln in lines if not ln.startswith("ok ")]
    assert not bad and r.returncode == 0, "\n".join(bad)

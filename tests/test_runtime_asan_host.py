"""CPU run of the library's host code under AddressSanitizer and, in a second
build, the undefined-behaviour sanitizer: tests/host/runtime_asan_host.cpp
replaces every HIP runtime function the host code calls with a host fake
(device memory is heap memory, copies and fills run at once) and stubs the
kernel launchers, then drives each host path twice in one process -- the
single-shot forward and adjoint with a snapshot spill, batched RTM with
f16 snapshots and the BF16 ring on step kernels 4 and 13, batched
modelling, the autotuner and the serving loop.  An upload, download, fill
or staging copy past the end of its buffer, a use after the pool freed a
buffer, a double destroy of a stream, event or graph, or a leak fails the
run.  What it checks is listed at the top of that file."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# host pass only: --cuda-host-only builds no GPU code, and every sanitizer
# flag goes to the host compiler
SAN = {
    "asan": ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"],
    "ubsan": ["-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
              "-fno-sanitize-recover=undefined"],
}


@pytest.mark.parametrize("san", ["asan", "ubsan"])
def test_host_paths_twice_under_sanitizer(tmp_path, san):
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "runtime_host")
    csrc = os.path.join(ROOT, "mnad-modeling-rtm_amd", "csrc")
    cmd = [HIPCC, "-std=c++17", "-O1", "-g", "--cuda-host-only",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           *SAN[san], "-I", os.path.join(ROOT, "include"), "-I", csrc,
           os.path.join(ROOT, "tests", "host", "runtime_asan_host.cpp"),
           *(os.path.join(csrc, f) for f in ("mnad_runtime.hip", "mnad_capi.hip",
                                             "mnad_autotune.hip", "mnad_server.hip")),
           "-o", exe, "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    lines = r.stdout.strip().splitlines()
    bad = [ln for ln in lines if ln.startswith("FAIL")]
    assert r.returncode == 0 and not bad, "\n".join(bad) + r.stderr[-4000:]
    assert sum(ln.startswith("ok ") for ln in lines) >= 60, r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, \
        r.stderr[-4000:]

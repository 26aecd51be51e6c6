"""CPU-side checks of the C-ABI library: it loads, exports exactly what
include/mnad.h declares, fails loudly without a GPU, and its host logic
(snapshot schedule, argument validation) matches the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "mnad.h")).read()
    return sorted(set(re.findall(r"MNAD_API[^;(]*?\b(mnad_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(mnad):
    names = _declared()
    assert len(names) >= 15
    assert sorted(mnad.EXPORTS) == names
    L = mnad.lib()
    for n in names:
        assert hasattr(L, n), n
    out = os.popen(f"nm -D --defined-only {mnad.LIB_PATH}").read()
    exported = sorted(set(re.findall(r" T (mnad_\w+)", out)))
    assert exported == names       # nothing else leaks out of the .so
    assert "gfx950" in mnad.version()


def test_library_contains_gfx950_code_object(mnad):
    blob = open(mnad.LIB_PATH, "rb").read()
    assert b"gfx950" in blob and b"k_step" in blob


def test_no_gpu_fails_loudly(mnad):
    if mnad.lib().mnad_device_count() > 0:
        pytest.skip("a GPU is visible here")
    with pytest.raises(mnad.MnadError) as e:
        mnad.Context(0)
    assert e.value.code == mnad.ERR_NOGPU


@pytest.mark.parametrize("nt,vmin,vmax", [(1000, 1500, 2500), (6000, 1500, 4500),
                                          (12000, 1500, 4500), (1922, 1500, 4440),
                                          (187, 1600, 2400), (4000, 1425, 4200)])
def test_snap_schedule_matches_oracle(mnad, oracle, nt, vmin, vmax):
    assert mnad.snap_schedule(nt, vmin, vmax) == oracle.snap_schedule(nt, vmin, vmax)


def test_snap_schedule_rejects_degenerate(mnad):
    with pytest.raises(mnad.MnadError) as e:
        mnad.snap_schedule(5, 1500, 1500)
    assert e.value.code == mnad.ERR_ARG


def test_matlab_mirror_validates_before_device(mnad):
    v = np.full((20, 20), 1500.0)
    w = np.zeros(10)
    with pytest.raises(mnad.MnadError):       # non-integer grid index
        mnad.fwd_PDE_exact_MNAD(10.0, 3.5, [2], 10.0, 3, [2], v, 1e-3, 10, w, 4, 1)
    with pytest.raises(mnad.MnadError):       # length(w) != nt
        mnad.fwd_PDE_exact_MNAD(10.0, 3, [2], 10.0, 3, [2], v, 1e-3, 11, w, 4, 1)
    with pytest.raises(mnad.MnadError):       # data rows != nt
        mnad.Born_adj_MNAD(10.0, 10.0, 12, 1e-3, v, [2], [2], np.zeros((10, 1)), 4, 1)

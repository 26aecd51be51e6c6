"""The MEX gateway (against the mock mx API) on the GPU: the reference's call
sequence fwd(indx=1) -> adj, including the cross-worker hand-off through a
snapshot file ('save' in one worker, 'load' in another)."""
import os

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mex(mnad):
    from mexgw import Mex
    return Mex()


def test_mex_forward_and_rtm_with_file_handoff(mex, tmp_path):
    print("mex fwd + adj with hand-off", flush=True)
    f = golden("rtm_rect")
    dx, dz, dt, d = float(f["dx"]), float(f["dz"]), float(f["dt"]), int(f["npml"])
    w = f["w"]
    nt = w.size
    xs, zs, xr, zr = int(f["xs0"]) + 1, int(f["zs0"]) + 1, f["xr0"] + 1, f["zr0"] + 1
    data = mex("fwd", dx, dz, dt, d, f["v"], w, nt, xs, zs, xr, zr, 3, 0)
    assert rel_l2(data, f["data0"]) < 1e-4
    os.environ["MNAD_SNAP"] = "f32"
    try:
        datas, illum = mex("fwd", dx, dz, dt, d, f["vs"], w, nt, xs, zs, xr, zr, 3, 1, nlhs=2)
    finally:
        del os.environ["MNAD_SNAP"]
    assert rel_l2(datas, f["datas0"]) < 1e-4
    assert rel_l2(illum, f["illum0"]) < 1e-4
    path = str(tmp_path / "UU_3.mnad")
    mex("save", 3, path)
    mex("release", 3)                      # as if the adjoint ran in another worker
    mex("load", 3, path)
    Im = mex("adj", dx, dz, nt, dt, f["vs"], xr, zr, data - datas, d, 3)
    assert Im.shape == (f["Im0"].size, 1)
    assert rel_l2(Im.ravel(), f["Im0"]) < 1e-4

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


def _two_shots(f):
    xs = [int(f["xs0"]) + 1, int(f["xs1"]) + 1]
    zs = [int(f["zs0"]) + 1, int(f["zs1"]) + 1]
    xr = np.concatenate([f["xr0"], f["xr1"]]) + 1
    zr = np.concatenate([f["zr0"], f["zr1"]]) + 1
    return xs, zs, xr, zr, [f["xr0"].size, f["xr1"].size]


def test_mex_batched_shots(mex, mnad, ctx):
    """'fwd_shots' and 'rtm_shots' (the drivers' parfor loops as one call):
    each shot's gathers meet the FP64 fixtures, the stacked image
    img = mean(Im)./mean(illum) meets the fixture image, and equals the image
    of the same call through the Python mirror bitwise (same library path)."""
    print("mex batched shots", flush=True)
    f = golden("rtm_two_shot")
    dx, dz, dt, d = float(f["dx"]), float(f["dz"]), float(f["dt"]), int(f["npml"])
    w = f["w"]
    nt = w.size
    xs, zs, xr, zr, nrec = _two_shots(f)
    D = mex("fwd_shots", dx, dz, dt, d, f["v"], w, nt, xs, zs, xr, zr, nrec, 0)
    assert D.shape == (nt, sum(nrec))
    assert rel_l2(D[:, :nrec[0]], f["data0"]) < 1e-4 and rel_l2(D[:, nrec[0]:], f["data1"]) < 1e-4
    img, ill, dobs, dsyn = mex("rtm_shots", dx, dz, dt, d, f["v"], f["vs"], w, nt, xs, zs, xr, zr,
                               nrec, 0, nlhs=4)
    assert rel_l2(img, f["img"]) < 1e-3
    assert rel_l2(dobs[:, :nrec[0]], f["data0"]) < 1e-4
    assert rel_l2(dsyn[:, nrec[0]:], f["datas1"]) < 1e-4
    g = mnad.Grid(f["vs"].shape[0], f["vs"].shape[1], d, dx, dz, dt)
    sid = ctx.stack_create(g)
    ctx.rtm_shots(sid, g, f["v"], f["vs"], w, [z - 1 for z in zs], [x - 1 for x in xs],
                  [(f["zr0"], f["xr0"]), (f["zr1"], f["xr1"])], snap_fmt=mnad.SNAP_F16)
    img_py, ill_py, _ = ctx.stack_image(sid, g)
    ctx.stack_destroy(sid)
    assert np.array_equal(img, img_py) and np.array_equal(ill, ill_py)


def test_mex_serving(mex):
    """'serve_start' / 'submit' / 'wait' / 'cancel' / 'serve_stop': a modelling
    and an RTM request served together return the fixture gather and image;
    waiting twice on a ticket, or on one never submitted, is mnad:state."""
    from mexgw import MexError
    print("mex serving", flush=True)
    fm, fr = golden("fwd_rect"), golden("rtm_rect")
    mex("serve_start", 0, 8, 0, 16)
    try:
        dx, dz, dt, d = float(fm["dx"]), float(fm["dz"]), float(fm["dt"]), int(fm["npml"])
        tm = mex("submit", 0, dx, dz, dt, d, fm["v"], np.zeros((0, 0)), fm["w"], fm["w"].size,
                 int(fm["xs"]) + 1, int(fm["zs"]) + 1, fm["xr"] + 1, fm["zr"] + 1, 1)
        dx, dz, dt, d = float(fr["dx"]), float(fr["dz"]), float(fr["dt"]), int(fr["npml"])
        tr = mex("submit", 1, dx, dz, dt, d, fr["v"], fr["vs"], fr["w"], fr["w"].size,
                 int(fr["xs0"]) + 1, int(fr["zs0"]) + 1, fr["xr0"] + 1, fr["zr0"] + 1, 1)
        gat, st, times = mex("wait", tm, -1, nlhs=3)
        assert st == 0 and times.shape == (1, 3) and times[0, 0] <= times[0, 1] <= times[0, 2]
        assert rel_l2(gat, fm["gather"]) < 1e-4
        im, st = mex("wait", tr, -1, nlhs=2)
        assert rel_l2(im.ravel(order="F"), fr["Im0"]) < 1e-3   # column-major, as Im
        for t in (tr, 12345):
            with pytest.raises(MexError) as e:
                mex("wait", t, 0)
            assert e.value.ident == "mnad:state"
    finally:
        mex("serve_stop")

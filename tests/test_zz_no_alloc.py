"""mnad_fwd_noalloc (the forward shot with every device buffer in the
library's fixed scratch array) against mnad_fwd, through the C-ABI, on two
calls in a row, the second on a smaller grid: gathers and final fields must
be bitwise equal (same kernels, same arithmetic, plain launches against graph
chunks, which are bitwise equal).  The context's memory pool must not change
across the allocation-free calls, and a shot larger than the array is
refused with MNAD_ERR_ARG.  Run last: it must also pass after the rest of the
suite has used the device."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _c1(nt):
    nzp = nxp = 201
    npml = 20
    V = np.where((np.arange(nzp) * 10.0)[:, None] < 1000.0, 1500.0, 2500.0) * np.ones((1, nxp))
    v = np.pad(V, npml, mode="edge")
    t = np.arange(nt) * 1e-3
    tau = np.pi * (t - 1.1 / 15.0) * 15.0
    w = (1 - 2 * tau**2) * np.exp(-tau**2)
    xr = npml + np.arange(nxp)
    zr = np.full(nxp, npml + 1)
    return v, w, npml, npml + 100, zr, xr


def test_noalloc_forward_equals_mnad_fwd(mnad):
    _say("allocation-free forward vs mnad_fwd")
    ctx = mnad.Context(0)
    try:
        v, w, npml, xs, zr, xr = _c1(700)
        g = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, 1e-3)
        f = golden("fwd_rect")
        gs = mnad.Grid(f["v"].shape[0], f["v"].shape[1], int(f["npml"]), float(f["dx"]),
                       float(f["dz"]), float(f["dt"]))
        cases = [(g, v, w, npml, xs, zr, xr),   # C1 size, then a smaller grid
                 (gs, f["v"], f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"])]
        for k, (gg, vv, ww, zs, xx, zz, rr) in enumerate(cases):
            pool0 = ctx.mem_info()
            got = ctx.fwd_noalloc(gg, vv, ww, zs, xx, zz, rr)
            assert ctx.mem_info() == pool0, "the allocation-free call touched the memory pool"
            ref = ctx.fwd(gg, vv, ww, zs, xx, zz, rr)
            for key in ("gather", "U", "V", "W"):
                assert np.array_equal(got[key], ref[key]), (k, key)
            _say(f"  call {k + 1} ({gg.nz}x{gg.nx}, nt {ww.size}): bitwise equal")
        if "gather" in f:   # the second call also meets the FP64 fixture
            rel = np.linalg.norm(got["gather"] - f["gather"]) / np.linalg.norm(f["gather"])
            assert rel < 1e-4
        # beyond the scratch array: a clear MNAD_ERR_ARG, nothing run
        big = np.full((1201, 2401), 2000.0, order="F")
        gb = mnad.Grid(1201, 2401, 30, 10.0, 10.0, 1e-3)
        with pytest.raises(mnad.MnadError) as ei:
            ctx.fwd_noalloc(gb, big, np.zeros(100), 30, 1200, [31], [1200])
        assert ei.value.code == mnad.ERR_ARG and "scratch" in str(ei.value)
    finally:
        ctx.close()

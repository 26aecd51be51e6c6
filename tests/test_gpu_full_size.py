"""The C3 (811 x 3461 padded, dx 10 m) and C5 (2061 x 8061 padded, dx 12.5 m)
grids at full size, with fewer time steps than their workloads (nt 6000 and
12000): the FP64 oracle cannot finish these in a test, so the checks are the
properties that hold at any size.

- Shots batched in one launch sequence give the gathers of the same shots run
  one per launch, bitwise; an RTM stack built from one batched call equals the
  stack built one shot per call, bitwise (each shot's image is added in shot
  order by the same compensated sum).
- The step kernels picked for the grids (two rows per wavefront at C3; at C5
  four columns per lane, and the two-step register-queue march the
  autotuner picks there, with the BF16 boundary ring in its RTM) agree with
  the default kernel to FP32 rounding (1e-5 relative L2 on gathers; the
  gather tolerance against FP64 is 1e-4).

Three shots per grid: on the first physical column (next to the left CPML
strip), at the centre and on the last column, with
split-spread receivers clipped at the model edges (ragged receiver counts)."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [("c3", (11, 2, 1), 500), ("c5", (6, 1, 4), 400), ("c5", (13, 2, 4), 400)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _workload(mnad, name, nt):
    import bench
    wl = bench.c3_workload(nt, 3) if name == "c3" else bench.c5_workload(nt, 3)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    return wl, g


@pytest.mark.parametrize("name,tun,nt", CASES)
def test_full_size_forward_batched_alone_and_default(mnad, name, tun, nt):
    _say(f"{name} full-size forward, kernel {tun} batched vs alone vs default")
    wl, g = _workload(mnad, name, nt)
    v, w, zs, xs, recs = wl["v"], wl["w"], wl["zs"], wl["xs"], wl["recs"]
    ns = len(xs)
    ctx = mnad.Context(0)
    try:
        ref = ctx.fwd_shots(g, v, w, zs, xs, recs, batch=1)
        ctx.set_tuning(kernel=tun[0], rows_per_thread=tun[1], block_rows=tun[2])
        bat = ctx.fwd_shots(g, v, w, zs, xs, recs, batch=ns)
        for k in range(ns):
            alone = ctx.fwd_shots(g, v, w, [zs[k]], [xs[k]], [recs[k]], batch=1)[0]
            assert bat[k].shape == (nt, recs[k][1].size)
            assert np.abs(ref[k]).max() > 0, k
            assert np.array_equal(bat[k], alone), k
            assert rel_l2(bat[k], ref[k]) < 1e-5, k
    finally:
        ctx.close()


@pytest.mark.parametrize("name,tun,nt", CASES)
def test_full_size_rtm_stack_batched_equals_one_shot_per_call(mnad, name, tun, nt):
    _say(f"{name} full-size RTM stack, kernel {tun}: one batched call vs one shot per call")
    wl, g = _workload(mnad, name, nt)
    v, vs, w, zs, xs, recs = wl["v"], wl["vs"], wl["w"], wl["zs"], wl["xs"], wl["recs"]
    ns = len(xs)
    snap = mnad.SNAP_RING_BF16 if tun[0] == 13 else mnad.SNAP_F16
    ctx = mnad.Context(0)
    try:
        ctx.set_tuning(kernel=tun[0], rows_per_thread=tun[1], block_rows=tun[2])
        sa = ctx.stack_create(g)
        ga = ctx.rtm_shots(sa, g, v, vs, w, zs, xs, recs, snap_fmt=snap, batch=ns,
                           want_gathers=True)
        img_a, ill_a, n_a = ctx.stack_image(sa, g)
        ctx.stack_destroy(sa)
        sb = ctx.stack_create(g)
        gb = ([], [])
        for k in range(ns):
            gk = ctx.rtm_shots(sb, g, v, vs, w, [zs[k]], [xs[k]], [recs[k]],
                               snap_fmt=snap, batch=1, want_gathers=True)
            gb[0].append(gk[0][0])
            gb[1].append(gk[1][0])
        img_b, ill_b, n_b = ctx.stack_image(sb, g)
        ctx.stack_destroy(sb)
    finally:
        ctx.close()
    assert n_a == n_b == ns
    for k in range(ns):
        assert np.abs(ga[0][k]).max() > 0, k
        assert np.array_equal(ga[0][k], gb[0][k]), k   # observed data, true model
        assert np.array_equal(ga[1][k], gb[1][k]), k   # synthetic data, migration model
    lit = ill_a > 0
    assert lit.mean() > 1e-3
    assert np.isfinite(img_a[lit]).all()
    assert np.array_equal(ill_a, ill_b)
    assert np.array_equal(img_a, img_b, equal_nan=True)   # img is 0/0 where nothing was lit

"""Serving loop (mnad_serve_* / mnad_submit / mnad_wait / mnad_cancel) on the
GPU: continuous batching of modelling and RTM requests of different grid
sizes.  Each result must equal the same request served alone (bitwise: a
shot's arithmetic does not depend on its neighbours in a launch) and meet the
FP64 fixtures at the stated tolerances (1e-4 gathers, 1e-3 f16 images)."""
import sys
import time

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grid(mnad, f, vkey="v"):
    v = f[vkey]
    return mnad.Grid(v.shape[0], v.shape[1], int(f["npml"]), float(f["dx"]), float(f["dz"]),
                     float(f["dt"]))


def _model_req(mnad, name):
    f = golden(name)
    return f, dict(kind=mnad.REQ_MODEL, grid=_grid(mnad, f), v=f["v"], w=f["w"],
                   zs=int(f["zs"]), xs=int(f["xs"]), zr=f["zr"], xr=f["xr"])


def _rtm_req(mnad, name, k=0, fmt=None):
    f = golden(name)
    return f, dict(kind=mnad.REQ_RTM, grid=_grid(mnad, f, "vs"), v=f["v"], v_smooth=f["vs"],
                   w=f["w"], zs=int(f[f"zs{k}"]), xs=int(f[f"xs{k}"]), zr=f[f"zr{k}"],
                   xr=f[f"xr{k}"], snap_fmt=mnad.SNAP_F16 if fmt is None else fmt)


@pytest.fixture()
def server(mnad):
    ctx = mnad.Context(0)
    ctx.serve_start(hbm_budget=0, max_active=64, chunk_steps=16, max_points=0)
    yield ctx
    ctx.serve_stop()
    ctx.close()


def _submit(ctx, r):
    kw = dict(r)
    return ctx.submit(kw.pop("kind"), kw.pop("grid"), kw.pop("v"), kw.pop("w"), kw.pop("zs"),
                      kw.pop("xs"), kw.pop("zr"), kw.pop("xr"), **kw)


def test_alone_equals_in_the_mix(mnad, server):
    _say("serving: alone vs mix")
    fm, rm = _model_req(mnad, "fwd_rect")
    fs, rs = _model_req(mnad, "fwd_square")
    fr, rr = _rtm_req(mnad, "rtm_two_shot", 0)
    fq, rq = _rtm_req(mnad, "rtm_rect", 0, fmt=mnad.SNAP_F32)
    alone = {}
    for key, r in (("m", rm), ("s", rs), ("r", rr), ("q", rq)):
        out, info = server.wait(_submit(server, r))
        assert info["status"] == mnad.OK
        alone[key] = out
    # the mix: everything in flight at once, plus a second copy of each
    tickets = {key: _submit(server, r) for key, r in (("m", rm), ("r", rr), ("s", rs), ("q", rq))}
    extra = [_submit(server, r) for r in (rq, rm, rr)]
    mixed = {key: server.wait(t)[0] for key, t in tickets.items()}
    for t in extra:
        server.wait(t)
    assert np.array_equal(mixed["m"]["gather"], alone["m"]["gather"])
    assert np.array_equal(mixed["s"]["gather"], alone["s"]["gather"])
    assert np.array_equal(mixed["r"]["image"], alone["r"]["image"])
    assert np.array_equal(mixed["q"]["image"], alone["q"]["image"])
    assert rel_l2(mixed["m"]["gather"], fm["gather"]) < 1e-4
    assert rel_l2(mixed["s"]["gather"], fs["gather"]) < 1e-4
    assert rel_l2(mixed["r"]["gather"], fr["data0"]) < 1e-4
    nzi, nxi = mixed["r"]["image"].shape
    assert rel_l2(mixed["r"]["image"].ravel(order="F"), fr["Im0"]) < 1e-3
    assert rel_l2(mixed["q"]["image"].ravel(order="F"), fq["Im0"]) < 1e-4


def test_late_arrival_joins_running_batch(mnad, server):
    """A short request submitted while a long one runs is admitted at the
    next chunk boundary and finishes first."""
    _say("serving: late arrival")
    f, long_r = _model_req(mnad, "fwd_square")
    long_r = dict(long_r)
    long_r["w"] = np.concatenate([f["w"]] * 40)          # 6400 steps
    _, short_r = _model_req(mnad, "fwd_rect")
    t_long = _submit(server, long_r)
    deadline = time.time() + 30
    while server.serve_status()["steps"] < 64 and time.time() < deadline:
        time.sleep(0.001)
    t_short = _submit(server, short_r)
    out_s, info_s = server.wait(t_short)
    out_l, info_l = server.wait(t_long)
    assert info_s["admit_s"] < info_l["done_s"]
    assert info_s["done_s"] < info_l["done_s"]


def test_budget_admission_cancel_and_rejection(mnad):
    _say("serving: budget, cancel, reject")
    f, r = _model_req(mnad, "fwd_square")
    ctx = mnad.Context(0)
    try:
        # find the reserved size of one request, then allow exactly one at a time
        ctx.serve_start(hbm_budget=0, chunk_steps=16)
        _, info = ctx.wait(_submit(ctx, r))
        one = info["bytes"]
        ctx.serve_stop()
        ctx.serve_start(hbm_budget=int(1.5 * one), chunk_steps=16)
        ts = [_submit(ctx, r) for _ in range(3)]
        t_cancel = _submit(ctx, r)
        ctx.cancel(t_cancel)
        infos = [ctx.wait(t)[1] for t in ts]
        for a, b in zip(infos, infos[1:]):          # one at a time, in order
            assert b["admit_s"] >= a["done_s"] - 1e-3
        with pytest.raises(mnad.MnadError) as e:
            ctx.wait(t_cancel)
        assert e.value.code == mnad.ERR_CANCELLED
        big = dict(r)
        big["w"] = np.concatenate([f["w"]] * 4000)   # gather alone exceeds the budget
        with pytest.raises(mnad.MnadError) as e:
            _submit(ctx, big)
        assert e.value.code == mnad.ERR_BUDGET
        bad = dict(r)
        bad["v"] = f["v"].copy()
        bad["v"][4, 4] = np.nan
        with pytest.raises(mnad.MnadError) as e:
            _submit(ctx, bad)
        assert e.value.code == mnad.ERR_ARG
        cfl = dict(r)
        cfl["grid"] = mnad.Grid(r["grid"].nz, r["grid"].nx, r["grid"].npml, 10.0, 10.0, 0.05)
        with pytest.raises(mnad.MnadError) as e:
            _submit(ctx, cfl)
        assert e.value.code == mnad.ERR_ARG
        ctx.serve_stop()
    finally:
        ctx.close()


def test_work_limit_admits_first_come_first_served(mnad):
    """max_points bounds the grid points stepping at once: with room for one
    request, requests run one after another in submission order."""
    _say("serving: work limit")
    f, r = _model_req(mnad, "fwd_square")
    pts = r["grid"].nz * r["grid"].nx
    ctx = mnad.Context(0)
    try:
        ctx.serve_start(max_points=pts, chunk_steps=16)
        infos = [ctx.wait(t)[1] for t in [_submit(ctx, r) for _ in range(3)]]
        for a, b in zip(infos, infos[1:]):
            assert b["admit_s"] >= a["done_s"] - 1e-3
        ctx.serve_stop()
    finally:
        ctx.close()


@pytest.mark.parametrize("kernel,rz,br", [(9, 2, 8), (7, 3, 16), (13, 4, 4)])
def test_server_on_whole_batch_or_two_step_tuning(mnad, kernel, rz, br):
    """A context tuned to a kernel that cannot step one level per launch (the
    resident-block kernel runs a whole batch per launch, the two-step kernels
    two levels) still serves: the server falls back to the default kernel and
    its results equal those of a server on a default context bitwise."""
    _say(f"serving with context kernel {kernel}")
    f, r = _model_req(mnad, "fwd_square")
    out = []
    for tun in (None, (kernel, rz, br)):
        ctx = mnad.Context(0)
        try:
            if tun:
                ctx.set_tuning(kernel=tun[0], rows_per_thread=tun[1], block_rows=tun[2])
            ctx.serve_start(hbm_budget=0, chunk_steps=16)
            res, info = ctx.wait(_submit(ctx, r))
            assert info["status"] == mnad.OK
            out.append(np.asarray(res["gather"]))
            ctx.serve_stop()
        finally:
            ctx.close()
    assert rel_l2(out[0], f["gather"]) < 1e-4
    assert np.array_equal(out[0], out[1])


def test_stop_and_double_wait_while_waiting(mnad):
    """Waiters blocked in mnad_wait while mnad_serve_stop runs wake with the
    request's final status (done, or cancelled by the stop) and never touch
    freed server state; two waiters on one ticket get the result exactly
    once, the other MNAD_ERR_STATE."""
    import threading
    _say("serving: stop while waiting, double wait")
    ctx = mnad.Context(0)
    try:
        ctx.serve_start(hbm_budget=0, max_active=4, chunk_steps=16, max_points=0)
        f, r = _model_req(mnad, "fwd_rect")
        # two waiters on one ticket
        t = _submit(ctx, r)
        got = []

        def wait_one(tk):
            try:
                out, info = ctx.wait(tk)
                got.append(("ok", out))
            except mnad.MnadError as e:
                got.append(("err", e.code))

        ths = [threading.Thread(target=wait_one, args=(t,)) for _ in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join(60)
        assert sorted(k for k, _ in got) == ["err", "ok"], got
        assert [c for k, c in got if k == "err"] == [mnad.ERR_STATE]
        out = [o for k, o in got if k == "ok"][0]
        assert rel_l2(out["gather"], f["gather"]) < 1e-4
        # a queue longer than max_active, a waiter on the last ticket, then stop
        tickets = [_submit(ctx, r) for _ in range(12)]
        got.clear()
        th = threading.Thread(target=wait_one, args=(tickets[-1],))
        th.start()
        time.sleep(0.05)
        ctx.serve_stop()
        th.join(60)
        assert not th.is_alive()
        assert got and (got[0][0] == "ok" or got[0][1] == mnad.ERR_CANCELLED), got
        with pytest.raises(mnad.MnadError) as ei:
            ctx.wait(tickets[0])
        assert ei.value.code == mnad.ERR_STATE   # no server any more
    finally:
        ctx.close()


def test_churn_cancellations_and_changing_mixes(mnad):
    """Thirty-six requests of four grid sizes and lengths from 140 to 1200
    steps, all submitted at once to a server with 16-step chunks, every fifth
    cancelled right after it is submitted (queued or already running): the
    active set changes at most chunk boundaries, so the padded tile count of
    the chunk graphs keeps changing, and cancelled requests give their memory
    back while the next chunk is already queued.  Every request that is not
    cancelled equals the same request served alone, bitwise; a cancelled one
    reports MNAD_ERR_CANCELLED unless it had already finished."""
    _say("serving: churn with cancellations")
    base = [_model_req(mnad, "fwd_rect")[1], _model_req(mnad, "fwd_square")[1],
            _rtm_req(mnad, "rtm_two_shot", 0)[1], _rtm_req(mnad, "rtm_rect", 0)[1],
            _rtm_req(mnad, "rtm_two_shot", 1)[1]]
    reqs = []
    for k in range(36):
        r = dict(base[k % len(base)])
        r["w"] = np.concatenate([r["w"]] * (1 + k % 6))[: 140 + 30 * k]
        reqs.append(r)
    ctx = mnad.Context(0)
    try:
        ctx.serve_start(hbm_budget=0, max_active=64, chunk_steps=16, max_points=0)
        alone = [ctx.wait(_submit(ctx, r))[0] for r in reqs]
        tickets = []
        for k, r in enumerate(reqs):
            tickets.append(_submit(ctx, r))
            if k % 5 == 4:
                ctx.cancel(tickets[-1])
        n_cancelled = 0
        for k, t in enumerate(tickets):
            try:
                out, info = ctx.wait(t)
            except mnad.MnadError as e:
                assert k % 5 == 4 and e.code == mnad.ERR_CANCELLED, (k, e)
                n_cancelled += 1
                continue
            assert info["status"] == mnad.OK
            key = "image" if reqs[k]["kind"] == mnad.REQ_RTM else "gather"
            assert np.array_equal(out[key], alone[k][key]), k
        assert n_cancelled >= 1
        ctx.serve_stop()
    finally:
        ctx.close()

"""HIP path (libmnad through its C-ABI) against the FP64 oracle and the golden
fixtures.  Tolerances (relative L2 against FP64): gathers and final fields
1e-4, RTM images 1e-3 -- the FP32 path's targets; the snapshot storage
format changes only the image error, stated per format."""
import sys

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu

GATHER_TOL = 1e-4
IMAGE_TOL = 1e-3


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grid(mnad, f, vkey="v"):
    v = f[vkey]
    return mnad.Grid(v.shape[0], v.shape[1], int(f["npml"]), float(f["dx"]), float(f["dz"]),
                     float(f["dt"]))


@pytest.mark.parametrize("name", ["fwd_square", "fwd_rect"])
def test_golden_forward_matlab_signature(mnad, name):
    """fwd_PDE_exact_MNAD with the reference's argument order and 1-based
    indices; fwd_rect has its source inside the corner CPML strips and a
    duplicated receiver."""
    _say(f"golden forward {name}")
    f = golden(name)
    dx, dz, dt, d = float(f["dx"]), float(f["dz"]), float(f["dt"]), int(f["npml"])
    w = f["w"]
    data = mnad.fwd_PDE_exact_MNAD(dx, f["xs"] + 1, f["xr"] + 1, dz, f["zs"] + 1, f["zr"] + 1,
                                   f["v"], dt, w.size, w, d, 1, *([None] * 9))
    assert rel_l2(data[0], f["gather"]) < GATHER_TOL


@pytest.mark.parametrize("name", ["fwd_square", "fwd_rect"])
def test_golden_final_fields(mnad, ctx, name):
    _say(f"golden final fields {name}")
    f = golden(name)
    g = _grid(mnad, f)
    r = ctx.fwd(g, f["v"], f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"])
    for k in ("gather", "U", "V", "W"):
        assert rel_l2(r[k], f[k]) < GATHER_TOL, k


# BF16 storage of the source wavefield is the boundary ring (ring_bf16), held
# to the image bound like every other format.  BF16 *snapshots* (8 mantissa
# bits of d2u) are kept only as a comparison format: they measure 0.9-1.2e-3 on
# these fixtures and 2.2e-3 on the full-size C2 image, so they do not meet the
# bound, and their test states that error (2e-3 guard) instead.
@pytest.mark.parametrize("fmt,tol", [("f32", 1e-4), ("f16", IMAGE_TOL), ("ring_bf16", IMAGE_TOL),
                                     ("ring_f16", IMAGE_TOL), ("bf16", 2e-3)])
@pytest.mark.parametrize("name", ["rtm_two_shot", "rtm_rect"])
def test_golden_rtm_matlab_signature(mnad, name, fmt, tol):
    """fwd(indx=1) on the smooth model, scattered data, Born_adj_MNAD: the
    per-shot image against the sparse transliteration, for every storage of
    the source wavefield."""
    _say(f"golden rtm {name} {fmt}")
    f = golden(name)
    dx, dz, dt, d = float(f["dx"]), float(f["dz"]), float(f["dt"]), int(f["npml"])
    w = f["w"]
    nt = w.size
    for k in range(int(f["nshots"])):
        xs, zs, xr, zr = f[f"xs{k}"] + 1, f[f"zs{k}"] + 1, f[f"xr{k}"] + 1, f[f"zr{k}"] + 1
        data = mnad.fwd_PDE_exact_MNAD(dx, xs, xr, dz, zs, zr, f["v"], dt, nt, w, d, k + 1,
                                       *([None] * 9))
        datas = mnad.fwd_PDE_exact_MNAD(dx, xs, xr, dz, zs, zr, f["vs"], dt, nt, w, d, k + 1,
                                        *([None] * 9), indx=1, snap_fmt=mnad.SNAP_FORMATS[fmt])
        assert rel_l2(data[0], f[f"data{k}"]) < GATHER_TOL
        assert rel_l2(datas[0], f[f"datas{k}"]) < GATHER_TOL
        ill = mnad.illumination(k + 1)
        assert rel_l2(ill, f[f"illum{k}"]) < 1e-4
        Im = mnad.Born_adj_MNAD(dx, dz, nt, dt, f["vs"], xr, zr, data[0] - datas[0], d, k + 1,
                                *([None] * 9))
        err = rel_l2(Im, f[f"Im{k}"])
        _say(f"  shot {k} image rel L2 {err:.2e}")
        assert err < tol


def test_born_adj_without_source_wavefield_fails(mnad, ctx):
    f = golden("rtm_rect")
    g = _grid(mnad, f, "vs")
    with pytest.raises(mnad.MnadError) as e:
        ctx.born_adj(g, f["vs"], np.zeros((f["w"].size, 3)), [5, 5, 5], [5, 6, 7], ss=999)
    assert e.value.code == mnad.ERR_STATE


def test_golden_rtm_stack_batched(mnad, ctx):
    """mnad_rtm_shots (both shots in one batch) reproduces the stacked image
    img = mean(Im)./mean(illum) of main_Acoustic_RTM.m:237-262."""
    _say("golden rtm stack batched")
    f = golden("rtm_two_shot")
    g = _grid(mnad, f, "vs")
    sid = ctx.stack_create(g)
    n = int(f["nshots"])
    zs = [int(f[f"zs{k}"]) for k in range(n)]
    xs = [int(f[f"xs{k}"]) for k in range(n)]
    recs = [(f[f"zr{k}"], f[f"xr{k}"]) for k in range(n)]
    gat = ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], zs, xs, recs,
                        snap_fmt=mnad.SNAP_F32, want_gathers=True)
    img, ill, ns = ctx.stack_image(sid, g)
    ctx.stack_destroy(sid)
    assert ns == n
    for k in range(n):
        assert rel_l2(gat[0][k], f[f"data{k}"]) < GATHER_TOL
        assert rel_l2(gat[1][k], f[f"datas{k}"]) < GATHER_TOL
    assert rel_l2(img, f["img"]) < IMAGE_TOL


def _c1_inputs():
    """C1: 201 x 201, dx = dz = 10 m, 1500 m/s above z = 1000 m, 2500 below,
    npml 20, Ricker 15 Hz, dt = 1 ms, nt = 1000, source at the centre of the
    first physical row, 201 receivers on the second."""
    nzp = nxp = 201
    npml = 20
    V = np.where((np.arange(nzp) * 10.0)[:, None] < 1000.0, 1500.0, 2500.0) * np.ones((1, nxp))
    v = np.pad(V, npml, mode="edge")
    dt, nt, fc = 1e-3, 1000, 15.0
    t = np.arange(nt) * dt
    tau = np.pi * (t - 1.1 / fc) * fc
    w = (1 - 2 * tau**2) * np.exp(-tau**2)
    xr = npml + np.arange(nxp)
    zr = np.full(nxp, npml + 1)
    return v, w, npml, npml + 100, zr, xr, dt


def test_c1_forward_vs_oracle(mnad, ctx, oracle):
    _say("C1 forward vs oracle")
    v, w, npml, xs, zr, xr, dt = _c1_inputs()
    g = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, dt)
    r = ctx.fwd(g, v, w, npml, xs, zr, xr)
    go = oracle.grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, dt)
    o = oracle.forward(go, v, w, npml, xs, zr, xr)
    errs = {k: rel_l2(r[k], o[k]) for k in ("gather", "U", "V", "W")}
    _say(f"  C1 rel L2 {errs}")
    for k, e in errs.items():
        assert e < GATHER_TOL, (k, e)


def test_graph_and_plain_launches_agree_bitwise(mnad, ctx):
    f = golden("fwd_rect")
    g = _grid(mnad, f)
    a = ctx.fwd(g, f["v"], f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"])
    b = ctx.fwd(g, f["v"], f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"],
                flags=mnad.FLAG_NO_GRAPH)
    for k in ("gather", "U", "V", "W"):
        assert np.array_equal(a[k], b[k]), k


def test_batched_shots_equal_single_shots(mnad, ctx):
    """Shots batched in one launch sequence give bit-identical gathers to the
    same shots run alone (the batch axis replaces parfor)."""
    _say("batched vs alone")
    v, w, npml, xs, zr, xr, dt = _c1_inputs()
    w = w[:300]
    g = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, dt)
    srcs = [npml + 10, npml + 100, npml + 190]
    recs = [(zr, xr), (zr[:50], xr[:50]), (zr[::3], xr[::3])]
    batched = ctx.fwd_shots(g, v, w, [npml] * 3, srcs, recs, batch=3)
    for k in range(3):
        alone = ctx.fwd(g, v, w, npml, srcs[k], recs[k][0], recs[k][1], want_state=False)
        assert np.array_equal(batched[k], alone["gather"]), k


@pytest.mark.parametrize("kernel,rz,br", [(7, 3, 16), (7, 5, 8), (6, 1, 2), (9, 2, 8),
                                          (10, 1, 2), (11, 2, 1), (13, 1, 1),
                                          (13, 2, 4)])
def test_two_step_and_four_column_kernels_batched(mnad, kernel, rz, br):
    """Eight C1 shots in one launch sequence (several rounds of workgroups
    per launch, so blocks that finish early overlap blocks that have not
    started) equal the same shots run alone bitwise, and the default kernel to
    FP32 rounding.  A two-step kernel whose halo points read CPML memories an
    owner block already overwrote would break the first comparison."""
    _say(f"batched kernel={kernel} rz={rz} br={br}")
    v, w, npml, xs, zr, xr, dt = _c1_inputs()
    w = w[:300]
    g = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, dt)
    srcs = [npml + 5 + 24 * k for k in range(8)]
    recs = [(zr, xr)] * 8
    ctx = mnad.Context(0)
    try:
        ref = ctx.fwd_shots(g, v, w, [npml] * 8, srcs, recs, batch=8)
        ctx.set_tuning(kernel=kernel, rows_per_thread=rz, block_rows=br)
        batched = ctx.fwd_shots(g, v, w, [npml] * 8, srcs, recs, batch=8)
        for k in (0, 3, 7):
            alone = ctx.fwd_shots(g, v, w, [npml], [srcs[k]], [recs[k]], batch=1)[0]
            assert np.array_equal(batched[k], alone), k
        for k in range(8):
            assert rel_l2(batched[k], ref[k]) < 1e-5, k
    finally:
        ctx.close()


def test_invalid_inputs_raise(mnad, ctx):
    f = golden("fwd_square")
    g = _grid(mnad, f)
    v = f["v"].copy()
    with pytest.raises(mnad.MnadError):
        ctx.fwd(g, v, f["w"], 0, 10**6, [1], [1])            # source outside the grid
    bad = v.copy()
    bad[3, 3] = np.nan
    with pytest.raises(mnad.MnadError):
        ctx.fwd(g, bad, f["w"], 5, 5, [1], [1])              # NaN velocity
    g2 = mnad.Grid(g.nz, g.nx, g.npml, g.dx, g.dz, 0.05)
    with pytest.raises(mnad.MnadError):
        ctx.fwd(g2, v, f["w"], 5, 5, [1], [1])               # CFL violated
    with pytest.raises(mnad.MnadError):
        mnad.fwd_PDE_exact_MNAD(10.0, 3.5, [2], 10.0, 3, [2], v, 1e-3, f["w"].size, f["w"],
                                int(f["npml"]), 1)           # non-integer index


@pytest.mark.parametrize("kernel,rz,br", [(0, 1, 4), (0, 2, 4), (0, 4, 4), (0, 1, 8), (0, 1, 16),
                                          (4, 1, 1), (4, 1, 2), (4, 1, 4), (4, 1, 8),
                                          (6, 1, 1), (6, 1, 2), (6, 1, 4),
                                          (7, 3, 4), (7, 3, 8), (7, 3, 16), (7, 5, 4), (7, 5, 8),
                                          (7, 7, 8), (7, 7, 16), (10, 1, 1), (10, 1, 2),
                                          (10, 1, 4), (11, 2, 1), (13, 1, 1), (13, 1, 4), (13, 4, 2)])
@pytest.mark.parametrize("strip_first", [0, 1])
def test_launch_shapes_agree(mnad, ctx, oracle, kernel, rz, br, strip_first):
    """Every kernel family / rows-per-thread / dispatch-order variant of the
    step kernel meets the FP64 fixture and agrees with the default shape to
    FP32 rounding (variants compile to different FMA contractions: ~1e-6
    after 140 steps)."""
    _say(f"launch shape kernel={kernel} rz={rz} br={br} strip_first={strip_first}")
    f = golden("fwd_rect")
    g = _grid(mnad, f)
    base = ctx.fwd(g, f["v"], f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"])
    old = ctx.tuning()
    try:
        ctx.set_tuning(kernel=kernel, rows_per_thread=rz, block_rows=br, strip_first=strip_first)
        r = ctx.fwd(g, f["v"], f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"])
    finally:
        ctx.set_tuning(**old)
    for k in ("gather", "U", "V", "W"):
        assert rel_l2(r[k], f[k]) < GATHER_TOL, k
        assert rel_l2(r[k], base[k]) < 1e-5, k


@pytest.mark.parametrize("kernel,rz,br", [(4, 1, 4), (4, 1, 1), (6, 1, 4),
                                          (6, 1, 1), (7, 3, 4), (7, 3, 16), (7, 7, 8), (9, 1, 8),
                                          (9, 4, 8), (10, 1, 1), (10, 1, 4), (11, 2, 1), (13, 1, 1), (13, 4, 4)])
def test_kernel_families_rtm_and_multi_tile(mnad, kernel, rz, br):
    """Each step-kernel family through the whole RTM (snapshots, imaging,
    receiver injection; f32, f16 and bf16 snapshots and the bf16 boundary ring,
    which kernels without the ring code hand to kernel 4) against the stacked
    fixture image, and on C1 (241 x 241: several tile columns, so the left and right
    CPML strips sit in different tiles) against the default kernel."""
    _say(f"kernel family kernel={kernel} rz={rz} br={br}")
    f = golden("rtm_two_shot")
    g = _grid(mnad, f, "vs")
    n = int(f["nshots"])
    zs = [int(f[f"zs{k}"]) for k in range(n)]
    xs = [int(f[f"xs{k}"]) for k in range(n)]
    recs = [(f[f"zr{k}"], f[f"xr{k}"]) for k in range(n)]
    v, w, npml, xsc, zr, xr, dt = _c1_inputs()
    w = w[:400]
    gc = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, dt)
    ctx = mnad.Context(0)   # fresh: no autotuned pick for these grids overrides the tuning
    base = ctx.fwd(gc, v, w, npml, xsc, zr, xr)
    try:
        ctx.set_tuning(kernel=kernel, rows_per_thread=rz, block_rows=br)
        for fmt, tol in ((mnad.SNAP_F32, 1e-4), (mnad.SNAP_F16, IMAGE_TOL), (mnad.SNAP_BF16, IMAGE_TOL),
                         (mnad.SNAP_RING_BF16, IMAGE_TOL)):
            sid = ctx.stack_create(g)
            gat = ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], zs, xs, recs, snap_fmt=fmt,
                                want_gathers=True)
            img, ill, ns = ctx.stack_image(sid, g)
            ctx.stack_destroy(sid)
            for k in range(n):
                assert rel_l2(gat[0][k], f[f"data{k}"]) < GATHER_TOL
                assert rel_l2(gat[1][k], f[f"datas{k}"]) < GATHER_TOL
            err = rel_l2(img, f["img"])
            _say(f"  stacked image rel L2 {err:.2e} (fmt {fmt})")
            assert err < tol
        r = ctx.fwd(gc, v, w, npml, xsc, zr, xr)
    finally:
        ctx.close()
    for k in ("gather", "U", "V", "W"):
        assert rel_l2(r[k], base[k]) < 1e-5, k


def test_autotune_and_serial_paths(mnad, ctx):
    """The autotuner picks a supported variant; RTM through the tuned batched
    path and through the plain serial path (one propagation per launch
    sequence, no graph) both reproduce the stacked fixture image."""
    _say("autotune + serial")
    f = golden("rtm_two_shot")
    g = _grid(mnad, f, "vs")
    tuning, per_launch, us = ctx.autotune(g, f["vs"], max_shots=4)
    assert per_launch in (1, 2, 4) and us > 0
    assert tuning["kernel"] in (0, 3, 4, 6)
    n = int(f["nshots"])
    zs = [int(f[f"zs{k}"]) for k in range(n)]
    xs = [int(f[f"xs{k}"]) for k in range(n)]
    recs = [(f[f"zr{k}"], f[f"xr{k}"]) for k in range(n)]
    imgs = []
    for flags in (0, mnad.FLAG_SERIAL | mnad.FLAG_NO_GRAPH):
        sid = ctx.stack_create(g)
        ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], zs, xs, recs, snap_fmt=mnad.SNAP_F32,
                      flags=flags)
        img, ill, ns = ctx.stack_image(sid, g)
        ctx.stack_destroy(sid)
        assert ns == n
        assert rel_l2(img, f["img"]) < IMAGE_TOL
        imgs.append(img)
    assert rel_l2(imgs[0], imgs[1]) < 1e-4


def test_pipelined_rtm_groups_match_one_group(mnad, ctx):
    """Shots in groups of one or two run pipelined (the adjoints of group g
    share their launches with the forwards of group g + 1); with batch = 0 the
    tuned curve picks the in-order or the pipelined schedule.  The stacked
    image, the illumination and the gathers equal those of one group of all
    shots and of the plain serial path, bitwise."""
    _say("pipelined rtm groups")
    f = golden("rtm_two_shot")
    g = _grid(mnad, f, "vs")
    ctx.autotune(g, f["vs"], max_shots=4)
    n = int(f["nshots"])
    zs = [int(f[f"zs{k}"]) for k in range(n)] * 2
    xs = [int(f[f"xs{k}"]) for k in range(n)] * 2
    recs = [(f[f"zr{k}"], f[f"xr{k}"]) for k in range(n)] * 2
    out = []
    for batch, flags in ((len(zs), 0), (1, 0), (2, 0), (0, 0),
                         (1, mnad.FLAG_SERIAL | mnad.FLAG_NO_GRAPH)):
        sid = ctx.stack_create(g)
        gat = ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], zs, xs, recs,
                            snap_fmt=mnad.SNAP_F16, batch=batch, flags=flags, want_gathers=True)
        img, ill, ns = ctx.stack_image(sid, g)
        ctx.stack_destroy(sid)
        assert ns == len(zs)
        out.append((img, ill, gat))
    for img, ill, gat in out[1:]:
        assert np.array_equal(img, out[0][0])
        assert np.array_equal(ill, out[0][1])
        for a, b in zip(gat[0] + gat[1], out[0][2][0] + out[0][2][1]):
            assert np.array_equal(a, b)
    assert rel_l2(out[0][0], f["img"]) < IMAGE_TOL


def test_pipelined_rtm_full_size_c2(mnad, ctx):
    """At the full C2 size (561 x 1061, nt 300, 4 shots), pipelined groups of
    one shot and one group of four give the same stacked image bitwise."""
    _say("pipelined rtm, full-size C2")
    wl = _c2_inputs(300)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    xs = [wl["xs"] - 300, wl["xs"] - 100, wl["xs"] + 100, wl["xs"] + 300]
    zs = [wl["zs"]] * 4
    recs = [(wl["zr"], wl["xr"])] * 4
    imgs = []
    for batch in (4, 1):
        sid = ctx.stack_create(g)
        ctx.rtm_shots(sid, g, v, wl["vs"], wl["w"], zs, xs, recs, snap_fmt=mnad.SNAP_F16,
                      batch=batch)
        img, ill, ns = ctx.stack_image(sid, g)
        ctx.stack_destroy(sid)
        assert ns == 4
        imgs.append((img, ill))
    lit = imgs[0][1] > 0
    assert lit.mean() > 0.001
    assert np.array_equal(imgs[0][1], imgs[1][1])
    assert np.array_equal(imgs[0][0][lit], imgs[1][0][lit])


@pytest.mark.parametrize("kernel,rz,br", [(9, 2, 8), (11, 2, 1)])
def test_full_size_c2_rtm_bitwise_as_default_kernel(mnad, kernel, rz, br):
    """The resident-block kernel (one cooperative launch for every step, LDS
    blocks exchanging edge records) and the two-row kernel (shared row loads)
    evaluate every point's update exactly as the default kernel: gathers,
    image and illumination equal bitwise on a full-size C2 RTM shot
    (561 x 1061, nt 300: two forwards in one launch, then the adjoint)."""
    _say(f"kernel {kernel} vs default, full-size C2 RTM")
    wl = _c2_inputs(300)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    out = []
    for tun in ((4, 1, 1), (kernel, rz, br)):
        ctx = mnad.Context(0)
        try:
            ctx.set_tuning(kernel=tun[0], rows_per_thread=tun[1], block_rows=tun[2])
            sid = ctx.stack_create(g)
            gat = ctx.rtm_shots(sid, g, v, wl["vs"], wl["w"], [wl["zs"]], [wl["xs"]],
                                [(wl["zr"], wl["xr"])], snap_fmt=mnad.SNAP_F16, batch=1,
                                want_gathers=True)
            img, ill, _ = ctx.stack_image(sid, g)
            ctx.stack_destroy(sid)
            out.append((gat[0][0], gat[1][0], img, ill))
        finally:
            ctx.close()
    assert np.abs(out[0][0]).max() > 0
    assert (out[0][3] > 0).mean() > 0.001
    for a, b in zip(out[0], out[1]):   # the image is NaN where nothing was lit
        assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("rows", ["", "24", "40", "64"])
def test_resident_kernel_block_heights(mnad, rows, monkeypatch):
    """Blocks of 4 (the smallest, several block rows per shot) up to 64 rows
    (C1 in four block rows, the last one partial) and three shots per launch:
    gathers and final fields equal the default kernel's bitwise, and a second
    launch on the same context (edge records left by the first) too."""
    _say(f"resident kernel, block height {rows or 'auto'}")
    if rows:
        monkeypatch.setenv("MNAD_RESIDENT_ROWS", rows)
    v, w, npml, xs, zr, xr, dt = _c1_inputs()
    w = w[:250]
    g = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, dt)
    srcs = [npml + 3, xs, v.shape[1] - npml - 4]
    recs = [(zr, xr)] * 3
    ctx = mnad.Context(0)
    try:
        ref = ctx.fwd_shots(g, v, w, [npml] * 3, srcs, recs, batch=3)
        ref1 = ctx.fwd(g, v, w, npml, xs, zr, xr)
        ctx.set_tuning(kernel=9, rows_per_thread=1, block_rows=8)
        for rep in range(2):
            got = ctx.fwd_shots(g, v, w, [npml] * 3, srcs, recs, batch=3)
            for k in range(3):
                assert np.array_equal(got[k], ref[k]), (rep, k)
        got1 = ctx.fwd(g, v, w, npml, xs, zr, xr)
        for key in ("gather", "U", "V", "W"):
            assert np.array_equal(got1[key], ref1[key]), key
    finally:
        ctx.close()


def test_resident_kernel_give_up_path(mnad, monkeypatch):
    """With the poll bound forced to zero from the host (MNAD_RESIDENT_SPIN=0),
    a block whose neighbour record is not yet out gives up at once; the launch
    still drains (every block exits), the call reports the documented error or
    completes exactly, and the same context then runs the batch exactly with
    the normal bound."""
    _say("resident kernel, give-up path")
    v, w, npml, xs, zr, xr, dt = _c1_inputs()
    w = w[:200]
    g = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, dt)
    srcs = [npml + 3, xs, v.shape[1] - npml - 4]
    recs = [(zr, xr)] * 3
    ctx = mnad.Context(0)
    try:
        ref = ctx.fwd_shots(g, v, w, [npml] * 3, srcs, recs, batch=3)
        ctx.set_tuning(kernel=9, rows_per_thread=2, block_rows=8)
        monkeypatch.setenv("MNAD_RESIDENT_SPIN", "0")
        gave_up = 0
        for _ in range(3):
            try:
                got = ctx.fwd_shots(g, v, w, [npml] * 3, srcs, recs, batch=3)
            except mnad.MnadError as e:
                assert e.code == mnad.ERR_HIP
                assert "gave up waiting for its neighbours" in str(e)
                gave_up += 1
            else:
                for k in range(3):
                    assert np.array_equal(got[k], ref[k]), k
        _say(f"  gave up in {gave_up} of 3 launches")
        monkeypatch.delenv("MNAD_RESIDENT_SPIN")
        got = ctx.fwd_shots(g, v, w, [npml] * 3, srcs, recs, batch=3)
        for k in range(3):
            assert np.array_equal(got[k], ref[k]), k
    finally:
        ctx.close()


@pytest.mark.parametrize("fmt", ["SNAP_F32", "SNAP_F16", "SNAP_BF16"])
def test_snapshot_file_handoff_is_exact(mnad, ctx, tmp_path, fmt):
    """The cross-process hand-off of the source wavefield (mnad_snap_save in
    one context, mnad_snap_load in another, replacing ./Wavefields/UU_<ss>.mat)
    gives the adjoint bitwise the image it computes from the snapshots kept on
    the device, for every storage format; truncated or foreign files and a
    missing file are rejected."""
    _say(f"snapshot file hand-off {fmt}")
    f = golden("rtm_rect")
    g = _grid(mnad, f, "vs")
    sf = getattr(mnad, fmt)
    w, zs, xs, zr, xr = f["w"], int(f["zs0"]), int(f["xs0"]), f["zr0"], f["xr0"]
    obs = ctx.fwd(g, f["v"], w, zs, xs, zr, xr, want_state=False)["gather"]
    syn = ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=5, indx=1, snap_fmt=sf, want_state=False)
    im_dev = ctx.born_adj(g, f["vs"], obs - syn["gather"], zr, xr, ss=5)
    ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=5, indx=1, snap_fmt=sf, want_state=False)
    path = str(tmp_path / "UU_5.mnad")
    ctx.snap_save(5, path)
    ctx.release(5)
    other = mnad.Context(0)
    try:
        other.snap_load(5, path)
        im_file = other.born_adj(g, f["vs"], obs - syn["gather"], zr, xr, ss=5)
        data = open(path, "rb").read()
        bad = str(tmp_path / "bad.mnad")
        with open(bad, "wb") as fh:
            fh.write(data[: len(data) // 2])
        with pytest.raises(mnad.MnadError):
            other.snap_load(6, bad)
        with open(bad, "wb") as fh:
            fh.write(b"NOTSNAPS" + data[8:])
        with pytest.raises(mnad.MnadError):
            other.snap_load(6, bad)
        with pytest.raises(mnad.MnadError):
            other.snap_load(6, str(tmp_path / "missing.mnad"))
    finally:
        other.close()
    assert np.array_equal(im_dev, im_file)
    assert np.abs(im_dev).max() > 0


def test_model_cache_follows_the_values(mnad, ctx):
    """The context keeps the device arrays of recent velocity models, keyed by
    the grid and a digest of the values.  The same values reuse them, changed
    values (also in the same host buffer) rebuild them, more models than the
    cache holds evict the oldest: every call equals a fresh context's result
    bitwise."""
    _say("model cache")
    f = golden("fwd_rect")
    g = _grid(mnad, f)
    args = (f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"])
    scales = [1.0, 1.05, 0.97, 1.02, 0.99, 1.01]
    fresh = mnad.Context(0)
    try:
        want = {sc: fresh.fwd(g, f["v"] * sc, *args, want_state=False)["gather"]
                for sc in scales}
    finally:
        fresh.close()
    buf = np.asfortranarray(f["v"].copy())
    for sc in [1.0, 1.05, 1.0] + scales + [1.0, 1.05]:
        buf[...] = f["v"] * sc            # same host buffer, new values
        got = ctx.fwd(g, buf, *args, want_state=False)["gather"]
        assert np.array_equal(got, want[sc]), sc


def _c2_inputs(nt):
    import bench
    return bench.c2_workload(nt)


def test_full_size_c2_linearity_exact(mnad, ctx):
    """At the full C2 size (561 x 1061) the path is linear: doubling the
    wavelet doubles the gather and quadruples the illumination and the image
    Im (img = Im./illum unchanged).  Scaling by 2 is exact for normal FP32
    numbers, but ahead of the wavefront the field passes through the
    subnormal range (~1e-39, relative precision ~1e-6), where round(2x) !=
    2 round(x); those seeds grow with the front to ~2e-6 relative (measured),
    so the bound is 1e-5, ten times below the parity tolerance."""
    _say("C2 full-size linearity")
    wl = _c2_inputs(600)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    a = ctx.fwd(g, v, wl["w"], wl["zs"], wl["xs"], wl["zr"], wl["xr"], want_state=False)
    b = ctx.fwd(g, v, 2 * wl["w"], wl["zs"], wl["xs"], wl["zr"], wl["xr"], want_state=False)
    assert np.abs(a["gather"]).max() > 0
    assert rel_l2(b["gather"], 2 * a["gather"]) < 1e-5
    imgs = []
    for scale in (1.0, 2.0):
        sid = ctx.stack_create(g)
        ctx.rtm_shots(sid, g, v, wl["vs"], scale * wl["w"], [wl["zs"]], [wl["xs"]],
                      [(wl["zr"], wl["xr"])], snap_fmt=mnad.SNAP_F16)
        img, ill, _ = ctx.stack_image(sid, g)
        ctx.stack_destroy(sid)
        imgs.append((img, ill))
    # after 600 steps the source wavefield has not reached the whole interior:
    # there illum = 0 and img = Im./illum is 0/0, as in the reference
    lit = imgs[0][1] > 0
    assert lit.mean() > 0.01 and np.isfinite(imgs[0][0][lit]).all()
    assert rel_l2(imgs[1][1], 4 * imgs[0][1]) < 1e-5
    assert rel_l2(imgs[1][0][lit], imgs[0][0][lit]) < 1e-4


def test_full_size_mirror_symmetry(mnad, ctx):
    """Laterally homogeneous C2-size model, source on the centre column: the
    gather is mirror symmetric to FP32 rounding."""
    _say("C2 full-size symmetry")
    wl = _c2_inputs(800)
    v = np.tile(wl["v"][:, wl["v"].shape[1] // 2][:, None], (1, wl["v"].shape[1]))
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    xc = v.shape[1] // 2
    xr = np.arange(wl["npml"], v.shape[1] - wl["npml"])
    zr = np.full(xr.size, wl["npml"] + 1)
    r = ctx.fwd(g, v, wl["w"], wl["zs"], xc, zr, xr, want_state=False)
    G = r["gather"]
    assert rel_l2(G[:, ::-1], G) < 1e-6


def test_failed_adjoint_keeps_the_source_wavefield(mnad, ctx):
    """A Born_adj_MNAD call that fails after the kept snapshots were handed to
    it (here: scattered data with a NaN, rejected on the host) leaves them in
    place, as the reference's ./Wavefields/UU_<ss>.mat survives a failed
    adjoint; the retry with valid data gives bitwise the image of a call that
    never failed."""
    _say("failed adjoint keeps the kept snapshots")
    f = golden("rtm_rect")
    g = _grid(mnad, f, "vs")
    w, zs, xs, zr, xr = f["w"], int(f["zs0"]), int(f["xs0"]), f["zr0"], f["xr0"]
    obs = ctx.fwd(g, f["v"], w, zs, xs, zr, xr, want_state=False)["gather"]
    syn = ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=9, indx=1, snap_fmt=mnad.SNAP_F16,
                  want_state=False)
    data = obs - syn["gather"]
    ref = ctx.born_adj(g, f["vs"], data, zr, xr, ss=9)
    ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=9, indx=1, snap_fmt=mnad.SNAP_F16, want_state=False)
    bad = data.copy()
    bad[bad.shape[0] // 2, 0] = np.nan
    with pytest.raises(mnad.MnadError) as ei:
        ctx.born_adj(g, f["vs"], bad, zr, xr, ss=9)
    assert ei.value.code == mnad.ERR_ARG
    got = ctx.born_adj(g, f["vs"], data, zr, xr, ss=9)
    assert np.array_equal(got, ref)
    with pytest.raises(mnad.MnadError) as ei:   # consumed by the successful call
        ctx.born_adj(g, f["vs"], data, zr, xr, ss=9)
    assert ei.value.code == mnad.ERR_STATE

"""RTM with the saved source-wavefield boundary ring (MNAD_SNAP_RING_BF16 /
MNAD_SNAP_RING_F16) on the GPU.

The forward saves (U, Vs, Ws) on a two-cell ring around the interior every
step as 16-bit closed-loop predictor residuals, and the adjoint's launches run
a reverse reconstruction of the source wavefield inside the CPML beside the
receiver back-propagation; d2u for the imaging condition comes from that
reconstruction.  The image must meet the FP64 fixtures at 1e-3 (the images'
bound), and the forward's gathers and illumination at 1e-4 as with every
other storage format (they do not depend on it)."""
import sys

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu

GATHER_TOL = 1e-4
IMAGE_TOL = 1e-3
RINGS = ["ring_bf16", "ring_f16"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grid(mnad, f, vkey="v"):
    v = f[vkey]
    return mnad.Grid(v.shape[0], v.shape[1], int(f["npml"]), float(f["dx"]), float(f["dz"]),
                     float(f["dt"]))


def _shots(f):
    n = int(f["nshots"])
    zs = [int(f[f"zs{k}"]) for k in range(n)]
    xs = [int(f[f"xs{k}"]) for k in range(n)]
    recs = [(f[f"zr{k}"], f[f"xr{k}"]) for k in range(n)]
    return n, zs, xs, recs


@pytest.mark.parametrize("fmt", RINGS)
@pytest.mark.parametrize("name", ["rtm_two_shot", "rtm_rect"])
def test_golden_rtm_ring_per_shot(mnad, ctx, name, fmt):
    """fwd(indx=1) with ring storage, scattered data, Born_adj_MNAD: the
    per-shot image against the sparse transliteration (rtm_two_shot pairs
    with shift 1, rtm_rect with shift 0 and dx != dz), and against the same
    path with FP32 snapshots."""
    _say(f"ring per shot {name} {fmt}")
    f = golden(name)
    g = _grid(mnad, f, "vs")
    sf = mnad.SNAP_FORMATS[fmt]
    for k in range(int(f["nshots"])):
        w, zs, xs, zr, xr = f["w"], int(f[f"zs{k}"]), int(f[f"xs{k}"]), f[f"zr{k}"], f[f"xr{k}"]
        obs = ctx.fwd(g, f["v"], w, zs, xs, zr, xr, want_state=False)["gather"]
        syn = ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=3, indx=1, snap_fmt=sf, want_state=False)
        assert rel_l2(syn["gather"], f[f"datas{k}"]) < GATHER_TOL
        assert rel_l2(syn["illum"], f[f"illum{k}"]) < 1e-4
        im = ctx.born_adj(g, f["vs"], obs - syn["gather"], zr, xr, ss=3)
        ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=4, indx=1, snap_fmt=mnad.SNAP_F32,
                want_state=False)
        im32 = ctx.born_adj(g, f["vs"], obs - syn["gather"], zr, xr, ss=4)
        err, e32 = rel_l2(im, f[f"Im{k}"]), rel_l2(im, im32)
        _say(f"  shot {k}: image rel L2 {err:.2e} vs FP64, {e32:.2e} vs FP32 snapshots")
        assert err < IMAGE_TOL
        assert e32 < IMAGE_TOL


@pytest.mark.parametrize("fmt", RINGS)
def test_golden_rtm_ring_stack_batched(mnad, ctx, fmt):
    """mnad_rtm_shots with ring storage (both shots in one batch: two
    adjoints and two reconstructions share the reverse launches) reproduces
    the stacked fixture image of main_Acoustic_RTM.m:237-262, and equals the
    same shots run one per call bitwise."""
    _say(f"ring stack batched {fmt}")
    f = golden("rtm_two_shot")
    g = _grid(mnad, f, "vs")
    n, zs, xs, recs = _shots(f)
    sf = mnad.SNAP_FORMATS[fmt]
    sid = ctx.stack_create(g)
    gat = ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], zs, xs, recs, snap_fmt=sf, batch=2,
                        want_gathers=True)
    img, ill, ns = ctx.stack_image(sid, g)
    ctx.stack_destroy(sid)
    assert ns == n
    for k in range(n):
        assert rel_l2(gat[0][k], f[f"data{k}"]) < GATHER_TOL
        assert rel_l2(gat[1][k], f[f"datas{k}"]) < GATHER_TOL
    err = rel_l2(img, f["img"])
    _say(f"  stacked image rel L2 {err:.2e}")
    assert err < IMAGE_TOL
    sid = ctx.stack_create(g)
    for k in range(n):
        ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], [zs[k]], [xs[k]], [recs[k]], snap_fmt=sf,
                      batch=1)
    img1, ill1, _ = ctx.stack_image(sid, g)
    ctx.stack_destroy(sid)
    assert np.array_equal(img, img1)
    assert np.array_equal(ill, ill1)


def test_ring_schedules_agree(mnad):
    """Four shots in groups of two through the three launch schedules: the
    pipelined one (explicit batch: a group's forwards share launches with the
    previous group's adjoints and reconstructions), the plain path (one
    propagation per launch sequence, no graph; an adjoint and its
    reconstruction still share their launches) and the autotuned pick.  The
    first two run the same per-shot arithmetic, so their images are equal
    bitwise; every image meets the fixture."""
    _say("ring schedules")
    f = golden("rtm_two_shot")
    g = _grid(mnad, f, "vs")
    n, zs, xs, recs = _shots(f)
    zs, xs, recs = zs * 2, xs * 2, recs * 2
    ctx = mnad.Context(0)
    try:
        imgs = []
        for batch, fl in ((2, 0), (1, mnad.FLAG_SERIAL | mnad.FLAG_NO_GRAPH), (0, 0)):
            if batch == 0:
                ctx.autotune(g, f["vs"], max_shots=4)
            sid = ctx.stack_create(g)
            ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], zs, xs, recs,
                          snap_fmt=mnad.SNAP_RING_BF16, batch=batch, flags=fl)
            imgs.append(ctx.stack_image(sid, g)[0])
            ctx.stack_destroy(sid)
    finally:
        ctx.close()
    for im in imgs:
        assert rel_l2(im, f["img"]) < IMAGE_TOL
    assert np.array_equal(imgs[0], imgs[1])
    assert rel_l2(imgs[2], imgs[0]) < 1e-4


@pytest.mark.parametrize("kernel,rz,br", [(4, 1, 4), (6, 1, 1), (6, 1, 4), (10, 1, 2),
                                          (11, 2, 1), (13, 1, 1), (13, 2, 4)])
def test_ring_kernel_families(mnad, kernel, rz, br):
    """Every kernel family that codes the ring and runs the reconstruction
    (one point per lane, one set of CPML memories prefetched, four and two
    columns per lane in the interior, two rows per wavefront, and the
    two-step register-queue kernel, whose ring codes, ring decodes and d2u
    come from the passes around its launches) reproduces the fixture image and
    agrees with the default kernel; on C1 (241 x 241, so the multi-column
    interior spans several tiles) as well."""
    _say(f"ring kernel {kernel},{rz},{br}")
    f = golden("rtm_two_shot")
    g = _grid(mnad, f, "vs")
    n, zs, xs, recs = _shots(f)
    v, w, npml, xsc = _c1()
    gc = mnad.Grid(v.shape[0], v.shape[1], npml, 10.0, 10.0, 1e-3)
    xr = npml + np.arange(201)
    zr = np.full(201, npml + 1)
    out = {}
    for tune in ((4, 1, 1), (kernel, rz, br)):
        ctx = mnad.Context(0)
        try:
            ctx.set_tuning(kernel=tune[0], rows_per_thread=tune[1], block_rows=tune[2])
            sid = ctx.stack_create(g)
            ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], zs, xs, recs,
                          snap_fmt=mnad.SNAP_RING_BF16, batch=2)
            a = ctx.stack_image(sid, g)[0]
            ctx.stack_destroy(sid)
            sid = ctx.stack_create(gc)
            ctx.rtm_shots(sid, gc, v, v * 0.98, w, [npml, npml], [xsc, xsc + 37], [(zr, xr)] * 2,
                          snap_fmt=mnad.SNAP_RING_BF16, batch=2)
            im, ill, _ = ctx.stack_image(sid, gc)
            ctx.stack_destroy(sid)
            # mean Im: where the wave has not arrived the illumination is 0 and img 0/0
            b = np.where(ill > 0, im * ill, 0.0)
            out[tune] = (a, b)
        finally:
            ctx.close()
    a4, b4 = out[(4, 1, 1)]
    a, b = out[(kernel, rz, br)]
    assert rel_l2(a, f["img"]) < IMAGE_TOL
    _say(f"  vs kernel 4: fixture grid {rel_l2(a, a4):.2e}, C1 {rel_l2(b, b4):.2e}")
    assert rel_l2(a, a4) < 1e-4
    assert rel_l2(b, b4) < 1e-4


def _c1():
    nzp = nxp = 201
    npml = 20
    V = np.where((np.arange(nzp) * 10.0)[:, None] < 1000.0, 1500.0, 2500.0) * np.ones((1, nxp))
    v = np.pad(V, npml, mode="edge")
    nt, fc = 600, 15.0
    t = np.arange(nt) * 1e-3
    tau = np.pi * (t - 1.1 / fc) * fc
    w = (1 - 2 * tau**2) * np.exp(-tau**2)
    return v, w, npml, npml + 100


@pytest.mark.parametrize("fmt", RINGS)
def test_ring_file_handoff_and_failed_adjoint(mnad, ctx, tmp_path, fmt):
    """The ring-stored source wavefield spilled to a file in one context and
    loaded in another (the UU_<ss>.mat hand-off between parfor workers) gives
    bitwise the image of the device-resident path; an adjoint that fails after
    taking it (a NaN in the data) leaves it in place, and the retry gives the
    same image bitwise; a truncated file is rejected."""
    _say(f"ring file hand-off {fmt}")
    f = golden("rtm_rect")
    g = _grid(mnad, f, "vs")
    sf = mnad.SNAP_FORMATS[fmt]
    w, zs, xs, zr, xr = f["w"], int(f["zs0"]), int(f["xs0"]), f["zr0"], f["xr0"]
    obs = ctx.fwd(g, f["v"], w, zs, xs, zr, xr, want_state=False)["gather"]
    syn = ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=5, indx=1, snap_fmt=sf, want_state=False)
    data = obs - syn["gather"]
    im_dev = ctx.born_adj(g, f["vs"], data, zr, xr, ss=5)
    ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=5, indx=1, snap_fmt=sf, want_state=False)
    bad = data.copy()
    bad[3, 2] = np.nan
    with pytest.raises(mnad.MnadError) as ei:
        ctx.born_adj(g, f["vs"], bad, zr, xr, ss=5)
    assert ei.value.code == mnad.ERR_ARG
    path = str(tmp_path / "UU_5.mnad")
    ctx.snap_save(5, path)
    assert np.array_equal(ctx.born_adj(g, f["vs"], data, zr, xr, ss=5), im_dev)
    other = mnad.Context(0)
    try:
        other.snap_load(5, path)
        im_file = other.born_adj(g, f["vs"], data, zr, xr, ss=5)
        raw = open(path, "rb").read()
        cut = str(tmp_path / "cut.mnad")
        with open(cut, "wb") as fh:
            fh.write(raw[: len(raw) // 2])
        with pytest.raises(mnad.MnadError):
            other.snap_load(6, cut)
    finally:
        other.close()
    assert np.array_equal(im_file, im_dev)
    assert rel_l2(im_dev, f["Im0"]) < IMAGE_TOL


def test_ring_file_handoff_two_step_kernel(mnad, tmp_path):
    """A ring forward kept on the two-step kernel (13, four time levels)
    images on that kernel, and spilled to a file and loaded in a context on
    the default kernel it images there too (the file holds the forward's last
    two levels as a two-level shot): the two images agree to 1e-4 and meet
    the fixture."""
    _say("ring file hand-off from kernel 13")
    f = golden("rtm_rect")
    g = _grid(mnad, f, "vs")
    w, zs, xs, zr, xr = f["w"], int(f["zs0"]), int(f["xs0"]), f["zr0"], f["xr0"]
    path = str(tmp_path / "UU_7.mnad")
    ctx = mnad.Context(0)
    other = mnad.Context(0)
    try:
        ctx.set_tuning(kernel=13, rows_per_thread=1, block_rows=1)
        obs = ctx.fwd(g, f["v"], w, zs, xs, zr, xr, want_state=False)["gather"]
        syn = ctx.fwd(g, f["vs"], w, zs, xs, zr, xr, ss=7, indx=1, snap_fmt=mnad.SNAP_RING_BF16,
                      want_state=False)
        data = obs - syn["gather"]
        ctx.snap_save(7, path)
        im13 = ctx.born_adj(g, f["vs"], data, zr, xr, ss=7)
        other.snap_load(7, path)
        im4 = other.born_adj(g, f["vs"], data, zr, xr, ss=7)
    finally:
        other.close()
        ctx.close()
    _say(f"  kernel 13 vs file on kernel 4: {rel_l2(im13, im4):.2e}")
    assert rel_l2(im13, im4) < 1e-4
    assert rel_l2(im13, f["Im0"]) < IMAGE_TOL
    assert rel_l2(im4, f["Im0"]) < IMAGE_TOL


def test_ring_needs_two_cpml_cells(mnad, ctx):
    """The ring sits on the two CPML cells next to the interior: a grid with
    npml < 2 is rejected with MNAD_ERR_ARG."""
    f = golden("rtm_rect")
    v = f["vs"][3:-3, 3:-3]   # npml 4 -> 1
    g = mnad.Grid(v.shape[0], v.shape[1], 1, float(f["dx"]), float(f["dz"]), float(f["dt"]))
    with pytest.raises(mnad.MnadError) as ei:
        ctx.fwd(g, v, f["w"], 2, 5, [2, 2], [4, 5], ss=1, indx=1, snap_fmt=mnad.SNAP_RING_BF16)
    assert ei.value.code == mnad.ERR_ARG


def test_kernel13_ring_sequences_replay_their_own_graph_chunks(mnad):
    """A kernel-13 ring RTM runs a forward sequence and then a sequence of
    adjoints and reconstructions whose padded tile tables can have the same
    length (here: a 48 x 64 grid, npml 8, three shots in one group, 16-row
    segments of four waves).  Each sequence must replay a graph chunk captured
    with its own passes (the reconstructions' ring decode and hand-off slot):
    with graph chunks of 8 launches the image equals the one without graphs
    bit for bit, and kernel 4's to 1e-4."""
    nz, nx, npml, nt = 48, 64, 8, 120
    _say(f"kernel-13 ring graph chunks {nz}x{nx}")
    j = np.arange(nz)[:, None] * np.ones((1, nx))
    v = np.asfortranarray(np.where(j < nz // 2, 1800.0, 2400.0))
    vs = np.asfortranarray(1800.0 + 600.0 * j / nz)
    w = np.zeros(nt)
    w[5] = 1.0
    g = mnad.Grid(nz, nx, npml, 10.0, 10.0, 1e-3)
    nr = nx - 2 * npml
    zs, xs = [npml] * 3, [npml + 2, nx // 2, nx - npml - 3]
    recs = []
    for k in range(3):
        xr = np.arange(npml, npml + nr, 1 + k)
        recs.append((np.full(xr.size, npml + 1), xr))
    images = {}
    for label, tune, flags in (("k13 graphs", (13, 2, 4), 0),
                               ("k13 no graph", (13, 2, 4), mnad.FLAG_NO_GRAPH),
                               ("k4 graphs", (4, 1, 1), 0)):
        ctx = mnad.Context(0)
        try:
            ctx.set_tuning(kernel=tune[0], rows_per_thread=tune[1], block_rows=tune[2],
                           graph_chunk=8)
            sid = ctx.stack_create(g)
            ctx.rtm_shots(sid, g, v, vs, w, zs, xs, recs, snap_fmt=mnad.SNAP_RING_BF16, batch=3,
                          flags=flags)
            img, ill, _ = ctx.stack_image(sid, g)
            ctx.stack_destroy(sid)
        finally:
            ctx.close()
        images[label] = np.where(ill > 0, img * ill, 0.0)
        assert np.isfinite(images[label]).all()
    e = rel_l2(images["k13 graphs"], images["k4 graphs"])
    _say(f"  kernel 13 vs kernel 4: {e:.2e}")
    assert np.array_equal(images["k13 graphs"], images["k13 no graph"])
    assert e < 1e-4


_MIN_CPML = {}


@pytest.mark.parametrize("npml", [2, 3])
@pytest.mark.parametrize("kernel,rz,br", [(4, 1, 1), (6, 1, 2), (10, 1, 1), (11, 2, 1), (13, 1, 1),
                                          (13, 2, 4)])
def test_ring_at_the_narrowest_cpml(mnad, oracle, npml, kernel, rz, br):
    """The ring takes the two CPML cells next to the interior, so with npml = 2
    it covers the whole CPML and its outer cells are the grid's edge (and with
    npml = 3 one CPML cell lies beyond it): a BF16-ring RTM with receivers on
    the surface and down a borehole near the source images within 1e-3 of the
    FP64 oracle on every kernel family that codes the ring."""
    _say(f"ring at npml {npml}, kernel {kernel},{rz},{br}")
    nz, nx, nt, dt = 100, 300, 300, 1e-3
    if npml not in _MIN_CPML:
        rng = np.random.default_rng(17)
        z = np.arange(nz)[:, None]
        v = np.clip(1800.0 + 700.0 * (z > 55) + 50.0 * rng.standard_normal((nz, nx)), 1500.0, 3000.0)
        vs = np.clip(v - 40.0, 1500.0, 3000.0)
        t = np.arange(nt) * dt
        tau = np.pi * (t - 1.1 / 25.0) * 25.0
        w = (1 - 2 * tau**2) * np.exp(-tau**2)
        xr = np.concatenate([np.arange(npml, nx - npml, 3), np.full(nz - 2 * npml - 2, 180)])
        zr = np.concatenate([np.full(xr.size - (nz - 2 * npml - 2), npml),
                             np.arange(npml + 1, nz - npml - 1)])
        go = oracle.grid(nz, nx, npml, 10.0, 10.0, dt)
        obs = oracle.forward(go, v, w, npml + 1, 150, zr, xr, want_state=False)["gather"]
        syn = oracle.forward(go, vs, w, npml + 1, 150, zr, xr, snap=True, want_state=False)
        im = oracle.adjoint(go, vs, zr, xr, obs - syn["gather"], syn["UU"])
        _MIN_CPML[npml] = dict(v=v, vs=vs, w=w, zr=zr, xr=xr, obs=obs,
                               im=im.reshape(nz - 2 * npml, nx - 2 * npml))
    c = _MIN_CPML[npml]
    g = mnad.Grid(nz, nx, npml, 10.0, 10.0, dt)
    ctx = mnad.Context(0)
    try:
        ctx.set_tuning(kernel=kernel, rows_per_thread=rz, block_rows=br)
        obs = ctx.fwd(g, c["v"], c["w"], npml + 1, 150, c["zr"], c["xr"], want_state=False)["gather"]
        syn = ctx.fwd(g, c["vs"], c["w"], npml + 1, 150, c["zr"], c["xr"], ss=4, indx=1,
                      snap_fmt=mnad.SNAP_RING_BF16, want_state=False)["gather"]
        im = ctx.born_adj(g, c["vs"], obs - syn, c["zr"], c["xr"], ss=4)
    finally:
        ctx.close()
    assert rel_l2(obs, c["obs"]) < GATHER_TOL
    e = rel_l2(im.reshape(nx - 2 * npml, nz - 2 * npml).T, c["im"])
    _say(f"  image vs FP64: {e:.2e}")
    assert e < IMAGE_TOL

"""Edge cases of the HIP path against the FP64 oracle (relative L2 <= 1e-4 on
gathers and final fields) for the default kernel, its CPML-prefetch variant,
the four-column kernel, the two-step kernel and the resident-block kernel: no CPML, the shortest runs, no
receivers, sources on the grid corners, receivers on every boundary
row/column, odd step counts for the two-step kernels (the last launch runs one
step), and a grid just large enough for its CPML strips."""
import sys

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4
KERNELS = [(4, 1, 1), (4, 1, 4), (6, 1, 2), (7, 3, 16), (9, 1, 8), (9, 4, 8),
           (10, 1, 1), (10, 1, 4), (11, 2, 1), (13, 1, 1), (13, 4, 4)]
IDS = ["k%d-%d-%d" % k for k in KERNELS]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _model(nz, nx, seed):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None]
    v = 1800.0 + 900.0 * (z > nz // 2) + 60.0 * rng.standard_normal((nz, nx))
    return np.clip(v, 1500.0, 3000.0)


def _wavelet(nt, dt, fc=25.0):
    t = np.arange(nt) * dt
    tau = np.pi * (t - 1.1 / fc) * fc
    return (1 - 2 * tau**2) * np.exp(-tau**2)


def _compare(mnad, oracle, kernel, v, npml, w, zs, xs, zr, xr, dx=10.0, dz=10.0, dt=1e-3):
    g = mnad.Grid(v.shape[0], v.shape[1], npml, dx, dz, dt)
    go = oracle.grid(v.shape[0], v.shape[1], npml, dx, dz, dt)
    o = oracle.forward(go, v, w, zs, xs, zr, xr)
    ctx = mnad.Context(0)
    try:
        k, rz, br = kernel
        ctx.set_tuning(kernel=k, rows_per_thread=rz, block_rows=br)
        r = ctx.fwd(g, v, w, zs, xs, zr, xr)
    finally:
        ctx.close()
    keys = ("U", "V", "W") + (("gather",) if len(zr) else ())
    for key in keys:
        ref = np.asarray(o[key], float)
        if np.linalg.norm(ref) > 0:
            e = rel_l2(r[key], ref)
        else:   # nothing has reached these points yet: both must be zero
            e = float(np.linalg.norm(np.asarray(r[key], float)))
        assert e < TOL, (key, e)
    return r


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
def test_no_cpml(mnad, oracle, kernel):
    _say(f"edge: npml 0, kernel {kernel}")
    v = _model(70, 300, 1)
    w = _wavelet(150, 1e-3)
    xr = np.arange(0, 300, 3)
    _compare(mnad, oracle, kernel, v, 0, w, 35, 150, np.full(xr.size, 2), xr)


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
def test_unequal_spacing_with_interior_tiles(mnad, oracle, kernel):
    """dx != dz on a grid large enough for every kernel's interior tiles (the
    fixtures with dx != dz are too small for them): kernel 13's interior waves
    then take the stencil weights from the shot instead of the dx == dz
    literals."""
    _say(f"edge: dx 12.5 m, dz 9 m, kernel {kernel}")
    v = _model(120, 400, 7)
    w = _wavelet(200, 1e-3)
    xr = np.arange(10, 390, 2)
    _compare(mnad, oracle, kernel, v, 10, w, 12, 200, np.full(xr.size, 11), xr, dx=12.5, dz=9.0)


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
@pytest.mark.parametrize("nt", [1, 2, 3])
def test_shortest_runs(mnad, oracle, kernel, nt):
    _say(f"edge: nt {nt}, kernel {kernel}")
    v = _model(40, 270, 2)
    w = _wavelet(40, 1e-3)[10:10 + nt]   # non-zero first samples
    # receivers far from the source and around it (the only ones a short run reaches)
    xr = np.concatenate([np.arange(5, 265, 7), np.arange(126, 135)])
    zr = np.concatenate([np.full(38, 6), np.full(9, 12)])
    _compare(mnad, oracle, kernel, v, 5, w, 12, 130, zr, xr)


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
def test_no_receivers(mnad, oracle, kernel):
    _say(f"edge: no receivers, kernel {kernel}")
    v = _model(40, 270, 3)
    w = _wavelet(90, 1e-3)
    r = _compare(mnad, oracle, kernel, v, 5, w, 12, 130, np.zeros(0, int), np.zeros(0, int))
    assert r["gather"].shape == (90, 0)


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
@pytest.mark.parametrize("corner", ["nw", "se"])
def test_source_on_grid_corner_and_boundary_receivers(mnad, oracle, kernel, corner):
    """Source at (0, 0) or (nz-1, nx-1), inside both CPML strips; receivers
    on the first and last rows and columns."""
    _say(f"edge: source corner {corner}, kernel {kernel}")
    nz, nx = 36, 280
    v = _model(nz, nx, 4)
    w = _wavelet(120, 1e-3)
    zs, xs = (0, 0) if corner == "nw" else (nz - 1, nx - 1)
    zr = np.concatenate([np.zeros(40, int), np.full(40, nz - 1), np.arange(nz), np.arange(nz)])
    xr = np.concatenate([np.arange(0, 280, 7), np.arange(3, 280, 7), np.zeros(nz, int),
                         np.full(nz, nx - 1)])
    _compare(mnad, oracle, kernel, v, 6, w, zs, xs, zr, xr)


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
def test_minimal_grid_for_cpml(mnad, oracle, kernel):
    """nz = nx = 2 npml + 1: a single interior point."""
    _say(f"edge: minimal grid, kernel {kernel}")
    v = _model(13, 13, 5)
    w = _wavelet(60, 1e-3)
    _compare(mnad, oracle, kernel, v, 6, w, 6, 6, [6, 0, 12], [6, 12, 0])


@pytest.mark.parametrize("kernel,rz,br", [(7, 3, 16), (7, 5, 8), (13, 1, 1), (13, 4, 4)])
def test_two_step_kernels_odd_step_count_rtm(mnad, kernel, rz, br):
    """rtm_rect has nt = 187: the last launch of a two-step kernel runs one step
    (and stores the level s+1 CPML memories itself)."""
    _say(f"edge: odd nt RTM, kernel {kernel}")
    f = golden("rtm_rect")
    assert f["w"].size % 2 == 1
    v = f["vs"]
    g = mnad.Grid(v.shape[0], v.shape[1], int(f["npml"]), float(f["dx"]), float(f["dz"]),
                  float(f["dt"]))
    ctx = mnad.Context(0)
    try:
        ctx.set_tuning(kernel=kernel, rows_per_thread=rz, block_rows=br)
        sid = ctx.stack_create(g)
        gat = ctx.rtm_shots(sid, g, f["v"], f["vs"], f["w"], [int(f["zs0"])], [int(f["xs0"])],
                            [(f["zr0"], f["xr0"])], snap_fmt=mnad.SNAP_F32, want_gathers=True)
        img, ill, ns = ctx.stack_image(sid, g)
        ctx.stack_destroy(sid)
    finally:
        ctx.close()
    assert rel_l2(gat[0][0], f["data0"]) < TOL
    assert rel_l2(gat[1][0], f["datas0"]) < TOL
    assert rel_l2(img, f["img"]) < 1e-3


_BOREHOLE = {}


def _borehole_case(oracle, zsrc):
    """120 x 400 grid (npml 10), source at row zsrc of column 200, receivers on
    the surface and down a borehole 300 m away (column 230) through the
    interior, so that the direct and reflected waves reach receivers within
    the 300 steps from every source depth; the FP64 oracle's observed and
    synthetic gathers and its image (computed once per zsrc)."""
    if zsrc not in _BOREHOLE:
        nz, nx, npml, nt, dt = 120, 400, 10, 300, 1e-3
        v = _model(nz, nx, 11)
        vs = np.clip(v - 40.0, 1500.0, 3000.0)
        w = _wavelet(nt, dt)
        xr = np.concatenate([np.arange(npml, nx - npml, 4), np.full(nz - 2 * npml - 4, 230)])
        zr = np.concatenate([np.full(xr.size - (nz - 2 * npml - 4), npml + 1),
                             np.arange(npml + 2, nz - npml - 2)])
        go = oracle.grid(nz, nx, npml, 10.0, 10.0, dt)
        obs = oracle.forward(go, v, w, zsrc, 200, zr, xr, want_state=False)["gather"]
        syn = oracle.forward(go, vs, w, zsrc, 200, zr, xr, snap=True, want_state=False)
        im = oracle.adjoint(go, vs, zr, xr, obs - syn["gather"], syn["UU"])
        _BOREHOLE[zsrc] = dict(nz=nz, nx=nx, npml=npml, dt=dt, v=v, vs=vs, w=w, zr=zr, xr=xr,
                               obs=obs, syn=syn["gather"],
                               im=im.reshape(nz - 2 * npml, nx - 2 * npml))
    return _BOREHOLE[zsrc]


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
@pytest.mark.parametrize("zsrc", [12, 64, 90])
def test_borehole_receivers_rtm(mnad, oracle, kernel, zsrc):
    """RTM with receivers down a borehole through the interior (the adjoint
    injects inside interior tiles at every depth, where kernel 13's interior
    waves take the march with receiver rows), the source near the surface or
    deep in the interior (kernel 13 then injects it on the interior path, and
    the reconstruction takes it back out there) or 20 rows above the bottom
    CPML: gathers at 1e-4 and the image at 1e-3 against the FP64 oracle, with
    f32 snapshots and the BF16 ring."""
    _say(f"edge: borehole receivers RTM, source row {zsrc}, kernel {kernel}")
    c = _borehole_case(oracle, zsrc)
    g = mnad.Grid(c["nz"], c["nx"], c["npml"], 10.0, 10.0, c["dt"])
    nzi, nxi = c["nz"] - 2 * c["npml"], c["nx"] - 2 * c["npml"]
    for fmt in (mnad.SNAP_F32, mnad.SNAP_RING_BF16):
        ctx = mnad.Context(0)
        try:
            k, rz, br = kernel
            ctx.set_tuning(kernel=k, rows_per_thread=rz, block_rows=br)
            obs = ctx.fwd(g, c["v"], c["w"], zsrc, 200, c["zr"], c["xr"],
                          want_state=False)["gather"]
            syn = ctx.fwd(g, c["vs"], c["w"], zsrc, 200, c["zr"], c["xr"], ss=2,
                          indx=1, snap_fmt=fmt, want_state=False)["gather"]
            im = ctx.born_adj(g, c["vs"], obs - syn, c["zr"], c["xr"], ss=2)
        finally:
            ctx.close()
        assert rel_l2(obs, c["obs"]) < TOL
        assert rel_l2(syn, c["syn"]) < TOL
        e = rel_l2(im.reshape(nxi, nzi).T, c["im"])
        _say(f"  format {fmt}: image vs FP64 {e:.2e}")
        assert e < 1e-3

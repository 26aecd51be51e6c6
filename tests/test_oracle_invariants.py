"""Closed-form and invariant checks that pin the oracle independently of
any transliteration: each is chosen so that a plausible mistake (a sign, an
offset, a transposed axis, a dropped CPML term) fails it."""
import numpy as np
import pytest


def _hom_grid(O, nz, nx, npml=0, dx=10.0, dz=10.0, dt=1e-3):
    return O.grid(nz, nx, npml, dx, dz, dt)


@pytest.mark.parametrize("dx,dz,theta", [(10.0, 10.0, 0.3), (12.0, 8.0, 1.1), (10.0, 10.0, 2.5)])
def test_plane_wave_one_step(oracle, dx, dz, theta):
    """u = cos(k.x - wt), V = -du/dx, W = -du/dz (x along columns, z along
    rows) solves the acoustic wave equation; one MNAD step must reproduce it
    to truncation error.  With V = +du/dx the residual is O(1)."""
    nz, nx, vel, dt = 40, 48, 2000.0, 1e-3
    g = _hom_grid(oracle, nz, nx, 0, dx, dz, dt)
    v = np.full((nz, nx), vel)
    kk = 2 * np.pi / 120.0
    kx, kz = kk * np.cos(theta), kk * np.sin(theta)
    om = vel * kk
    z, x = np.meshgrid(np.arange(nz) * dz, np.arange(nx) * dx, indexing="ij")

    def state(t, sgn=-1.0):
        ph = kx * x + kz * z - om * t
        return np.cos(ph), sgn * (-kx * np.sin(ph)), sgn * (-kz * np.sin(ph))

    sl = np.s_[4:-4, 4:-4]
    for sgn, expect_small in ((-1.0, True), (1.0, False)):
        p, c = oracle.steps(g, v, 1, state(-dt, sgn), state(0.0, sgn))
        ex = state(dt, sgn)
        res_u = np.abs(c[0] - ex[0])[sl].max()
        res_v = np.abs(c[1] - ex[1])[sl].max() / kk
        res_w = np.abs(c[2] - ex[2])[sl].max() / kk
        if expect_small:
            assert res_u < 1e-6 and res_v < 1e-4 and res_w < 1e-4, (res_u, res_v, res_w)
        else:
            assert max(res_u, res_v, res_w) > 1e-2


def test_auu_row_sum_and_avv_closed_form(oracle):
    """Auu 1 = 2 (Auu_gen.m: the a and b terms cancel); Avv 1 = 2 - 15 rx +
    15 rx^2 and Aww 1 = 2 - 15 rz + 15 rz^2 (rx = (v dt/dx)^2)."""
    nz, nx, dx, dz, dt = 12, 14, 11.0, 7.0, 1.3e-3
    g = _hom_grid(oracle, nz, nx, 0, dx, dz, dt)
    rng = np.random.default_rng(1)
    v = 1500 + 2000 * rng.random((nz, nx))
    one, zero = np.ones((nz, nx)), np.zeros((nz, nx))
    a = v**2 * dt**2
    rx, rz = a / dx**2, a / dz**2
    sl = np.s_[2:-2, 2:-2]
    oU, oV, oW = oracle.apply_A(g, v, one, zero, zero)
    assert np.abs(oU - 2)[sl].max() < 1e-13
    oU, oV, oW = oracle.apply_A(g, v, zero, one, zero)
    assert np.abs(oV - (2 - 15 * rx + 15 * rx**2))[sl].max() < 1e-13
    oU, oV, oW = oracle.apply_A(g, v, zero, zero, one)
    assert np.abs(oW - (2 - 15 * rz + 15 * rz**2))[sl].max() < 1e-13


@pytest.mark.parametrize("alpha,beta", [(0.3, 0.0), (0.0, -0.7), (1.1, 0.4)])
def test_static_linear_field_is_preserved(oracle, alpha, beta):
    """u = alpha x + beta z with V = -alpha, W = -beta is a static solution:
    A (U,V,W) = 2 (U,V,W) at every point whose stencil stays in the grid."""
    nz, nx, dx, dz, dt = 13, 17, 10.0, 6.0, 1e-3
    g = _hom_grid(oracle, nz, nx, 0, dx, dz, dt)
    rng = np.random.default_rng(2)
    v = 1500 + 1500 * rng.random((nz, nx))
    z, x = np.meshgrid(np.arange(nz) * dz, np.arange(nx) * dx, indexing="ij")
    U = alpha * x + beta * z
    V = np.full((nz, nx), -alpha)
    W = np.full((nz, nx), -beta)
    oU, oV, oW = oracle.apply_A(g, v, U, V, W)
    sl = np.s_[2:-2, 2:-2]
    assert np.abs(oU - 2 * U)[sl].max() < 1e-11 * max(1, np.abs(U).max())
    assert np.abs(oV - 2 * V)[sl].max() < 1e-12
    assert np.abs(oW - 2 * W)[sl].max() < 1e-12


def test_cpml_coefficients_off_strip_and_profile(oracle):
    """C_cof.m: c1 = 1, c2 = c3 = 0 off the side's strip (the memories stay
    exactly zero there); inside, damping grows toward the outer edge and the
    left/right profiles mirror each other for a mirrored model."""
    nz, nx, npml = 30, 40, 8
    g = _hom_grid(oracle, nz, nx, npml)
    v = np.tile(np.linspace(1500, 2500, nz)[:, None], (1, nx))
    masks = [np.s_[:, :npml], np.s_[:, nx - npml:], np.s_[:npml, :], np.s_[nz - npml:, :]]
    for side in range(4):
        c1, c2, c3 = oracle.cpml_coeffs(g, v, side)
        on = np.zeros((nz, nx), bool)
        on[masks[side]] = True
        assert np.all(c1[~on] == 1) and np.all(c2[~on] == 0) and np.all(c3[~on] == 0)
        assert np.all(c3[on] > 0) and np.all(c1[on] < 1)
    l1, l2, l3 = oracle.cpml_coeffs(g, v, 0)
    r1, r2, r3 = oracle.cpml_coeffs(g, v, 1)
    assert np.allclose(l1[:, :npml], r1[:, ::-1][:, :npml], rtol=0, atol=1e-15)
    assert np.allclose(l3[:, :npml], r3[:, ::-1][:, :npml], rtol=0, atol=1e-15)
    assert np.all(np.diff(l3[:, :npml], axis=1) < 0)        # strongest at the edge


@pytest.mark.parametrize("R", [1e-4, 1e-3])
def test_cpml_coefficient_values_from_the_profile(oracle, R):
    """C_cof.m's three coefficients at every strip point of every side are
    pinned by what the damping profile must satisfy, not by retyping it:
      - c1 = (1 - s)/(1 + s) and c3 = 2 s/(1 + s) with s = sigma dt / 2, so
        c1 + c3 = 1 exactly;
      - c2 / c3 = sigma' / sigma = 2 / x for the quadratic profile
        sigma = sigma0 x^2, with x the distance into the layer counted from
        the interior (one cell h at the first strip cell, D = npml h at the
        outer one, C_cof.m:6,38);
      - the profile absorbs to the target reflection R for a normal-incidence
        wave: (2 / v) sum_k sigma(x_k) h = ln(1/R) (2 n^3 + 3 n^2 + n) / (2 n^3),
        the discrete form of (2/v) int_0^D sigma dx = ln(1/R) with n = npml
        (sigma0 = 3 v ln(1/R) / (2 D^3)).
    A wrong power, scale, sign of x or an off-by-one strip index fails one."""
    nz, nx, npml = 26, 34, 7
    dx, dz, dt = 11.0, 9.0, 1e-3
    g = oracle.grid(nz, nx, npml, dx, dz, dt)
    rng = np.random.default_rng(3)
    v = 1500.0 + 2000.0 * rng.random((nz, nx))
    n = npml
    target = np.log(1 / R) * (2 * n**3 + 3 * n**2 + n) / (2 * n**3)
    for side in range(4):
        c1, c2, c3 = oracle.cpml_coeffs(g, v, side, R=R)
        h = dx if side < 2 else dz
        # distance into the layer of each strip cell, and the strip's cells
        if side == 0:
            x = (n - np.arange(n)) * h
            sl = lambda a: a[:, :n]                         # noqa: E731
        elif side == 1:
            x = (np.arange(n) + 1) * h
            sl = lambda a: a[:, nx - n:]                    # noqa: E731
        elif side == 2:
            x = (n - np.arange(n)) * h
            sl = lambda a: a[:n, :].T                       # noqa: E731
        else:
            x = (np.arange(n) + 1) * h
            sl = lambda a: a[nz - n:, :].T                  # noqa: E731
        a1, a2, a3 = sl(c1), sl(c2), sl(c3)
        assert np.allclose(a1 + a3, 1.0, rtol=0, atol=4e-16)
        assert np.allclose(a2 * x[None, :], 2.0 * a3, rtol=1e-13, atol=0)
        sigma = a3 / (dt * (1.0 - a3 / 2.0))                 # c3 = sigma dt / (1 + sigma dt / 2)
        vv = sl(v)
        assert np.allclose((2.0 / vv * sigma * h).sum(axis=1), target, rtol=1e-12)


def test_arrival_times_follow_the_velocity(oracle):
    """In a homogeneous 2000 m/s medium the wave reaches receivers 300, 600
    and 900 m from the source along x and along z with lags (r2 - r1) / v
    between them (cross-correlation peak of the traces, refined to a
    fraction of a sample): a wrong velocity scale, a swapped axis or a
    wrong dx/dz in the operators moves them.  Dispersion at 12 points per
    dominant wavelength keeps the error near 3e-4 (bound 2e-3)."""
    nz, nx, npml = 221, 221, 20
    dx, dt, v0, fc, nt = 10.0, 1e-3, 2000.0, 15.0, 800
    g = oracle.grid(nz, nx, npml, dx, dx, dt)
    v = np.full((nz, nx), v0)
    c = nz // 2
    offs = np.array([30, 60, 90])
    zr = np.concatenate([np.full(3, c), c + offs])
    xr = np.concatenate([c + offs, np.full(3, c)])
    r = oracle.forward(g, v, _ricker(fc, dt, nt), c, c, zr, xr, want_state=False)["gather"]

    def lag(a, b):
        cc = np.correlate(b, a, mode="full")
        k = int(np.argmax(cc))
        y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
        return (k - (a.size - 1)) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)

    for base in (0, 3):   # along x, along z
        for q in (1, 2):
            want = (offs[q] - offs[0]) * dx / v0 / dt
            got = lag(r[:, base], r[:, base + q])
            assert abs(got - want) < 2e-3 * want, (base, q, got, want)


def _ricker(fc, dt, nt):
    t = np.arange(nt) * dt
    tau = np.pi * (t - 1.1 / fc) * fc
    return (1 - 2 * tau**2) * np.exp(-tau**2)


def test_cpml_absorbs(oracle):
    """A pulse leaving a homogeneous box: the late-time trace near the source
    differs from a box large enough to have no reflection yet by far less with
    the CPML (npml 20) than with the bare zero boundary (npml 0)."""
    vel, dx, dt, nt = 2000.0, 10.0, 1e-3, 420
    w = _ricker(25.0, dt, nt)

    def trace(n, npml):
        g = oracle.grid(n, n, npml, dx, dx, dt)
        v = np.full((n, n), vel)
        c = n // 2
        return oracle.forward(g, v, w, c, c, [c + 3], [c + 3], want_state=False)["gather"][:, 0]

    ref = trace(221, 0)                 # reflections arrive after ~1.1 s
    with_pml = trace(101, 20)
    bare = trace(61, 0)
    late = slice(250, nt)
    e_ref = np.linalg.norm(ref[late]) + np.linalg.norm(ref)
    err_pml = np.linalg.norm(with_pml[late] - ref[late]) / e_ref
    err_bare = np.linalg.norm(bare[late] - ref[late]) / e_ref
    assert err_pml < 0.05 * err_bare, (err_pml, err_bare)
    assert err_pml < 0.02


def test_mirror_symmetry_and_linearity(oracle):
    """Laterally homogeneous model, centred source: the gather is mirror
    symmetric, U and W are even and V is odd in x; doubling w doubles all."""
    nz, nx, npml, dt, nt = 40, 51, 8, 1e-3, 150
    g = oracle.grid(nz, nx, npml, 10.0, 10.0, dt)
    v = np.tile(np.where(np.arange(nz) < 22, 1500.0, 2300.0)[:, None], (1, nx))
    w = _ricker(25.0, dt, nt)
    c = nx // 2
    xr = np.arange(npml, nx - npml)
    zr = np.full(xr.size, npml + 1)
    r = oracle.forward(g, v, w, npml, c, zr, xr)
    G = r["gather"]
    scale = np.abs(G).max()
    assert np.abs(G - G[:, ::-1]).max() < 1e-12 * scale
    assert np.abs(r["U"] - r["U"][:, ::-1]).max() < 1e-12 * np.abs(r["U"]).max()
    assert np.abs(r["V"] + r["V"][:, ::-1]).max() < 1e-12 * np.abs(r["V"]).max()
    assert np.abs(r["W"] - r["W"][:, ::-1]).max() < 1e-12 * np.abs(r["W"]).max()
    r2 = oracle.forward(g, v, 2 * w, npml, c, zr, xr)
    assert np.abs(r2["gather"] - 2 * G).max() < 1e-12 * scale


def test_time_reversibility(oracle):
    """Without CPML the leapfrog form X+ = A X - X- is exactly reversible:
    n steps forward, swap the levels, n steps back returns the start."""
    nz, nx = 36, 40
    g = oracle.grid(nz, nx, 0, 10.0, 10.0, 1e-3)
    rng = np.random.default_rng(3)
    v = 1500 + 1000 * rng.random((nz, nx))
    z, x = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    bump = lambda cz, cx, s: np.exp(-((z - cz)**2 + (x - cx)**2) / s)
    X0 = [bump(18, 20, 9), 0.01 * bump(17, 19, 8), -0.02 * bump(19, 21, 10)]
    X1 = [bump(18, 20, 9.5), 0.011 * bump(17, 19, 8), -0.019 * bump(19, 21, 10)]
    p, c = oracle.steps(g, v, 40, X0, X1)
    p2, c2 = oracle.steps(g, v, 40, c, p)
    for a, b in zip(c2, X0):
        assert np.abs(a - b).max() < 1e-9 * np.abs(b).max()
    for a, b in zip(p2, X1):
        assert np.abs(a - b).max() < 1e-9 * np.abs(b).max()

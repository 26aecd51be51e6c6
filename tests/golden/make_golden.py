"""Write the golden fixtures under tests/golden/ from a literal FP64
transliteration of the MATLAB reference.

This is an independent restatement, kept deliberately close to the MATLAB
statements (sparse operator assembly with MATLAB's column-major linear
indices, full-grid CPML memories, 1-based indices), so that agreement with the
stencil-form C oracle (oracle/mnad_oracle.c) pins both. It needs only numpy and
scipy and runs in seconds on tiny grids:

    python tests/golden/make_golden.py

Sources followed (paths relative to the reference root):
  SRC/MNAD/{Auu,...,Aww}_gen.m, SRC/MNAD/Ux123_gen.m, SRC/MNAD/Uz123_gen.m,
  SRC/MNAD/C_cof.m, SRC/MNAD/fwd_PDE_exact_MNAD.m,
  SRC/migration/Born_adj_MNAD.m (with U_r read as U and data passed as a
  matrix), SRC/migration/main_Acoustic_RTM.m:191-262 (smoothing, scattered
  data, illumination mean, image stack), SRC/MNAD/v_smooth.m,
  SRC/modeling/main_Acoustic_Data.m:134-145 (Ricker, edge padding).
"""
from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))


# --------------------------------------------------------------- operators
def _assemble(nz, nx, blocks):
    """sparse(I, J, vals, N, N) with MATLAB column-major linear indices.

    blocks: list of (jrange, irange, dj, di, vals) where j/i are 1-based
    ranges of the OUTPUT point (row index I = idx(j,i)), the input point is
    (j+dj, i+di) (column index J) and vals is evaluated on v(j,i)."""
    N = nz * nx
    idx = np.arange(N).reshape(nz, nx, order="F")  # idx(j,i), 0-based values
    I_all, J_all, V_all = [], [], []
    for (j0, j1), (i0, i1), dj, di, vals in blocks:
        jj = np.arange(j0, j1 + 1) - 1
        ii = np.arange(i0, i1 + 1) - 1
        I = idx[np.ix_(jj, ii)]
        J = idx[np.ix_(jj + dj, ii + di)]
        I_all.append(I.ravel(order="F"))
        J_all.append(J.ravel(order="F"))
        V_all.append(np.asarray(vals(jj, ii)).ravel(order="F"))
    return sp.csc_matrix((np.concatenate(V_all), (np.concatenate(I_all),
                          np.concatenate(J_all))), shape=(N, N))


def build_ops(dx, dz, dt, v):
    nz, nx = v.shape

    def A(jj, ii):
        return v[np.ix_(jj, ii)] ** 2 * dt ** 2

    def B(jj, ii):
        return v[np.ix_(jj, ii)] ** 4 * dt ** 4

    ops = {}
    # Auu_gen.m:25-86
    ops["Auu"] = _assemble(nz, nx, [
        ((2, nz), (1, nx), -1, 0, lambda j, i: (24 / 11) * A(j, i) / dz**2 - B(j, i) * (15 / (11 * dz**4) + 1 / (dx**2 * dz**2))),
        ((1, nz), (2, nx), 0, -1, lambda j, i: (24 / 11) * A(j, i) / dx**2 - B(j, i) * (15 / (11 * dx**4) + 1 / (dx**2 * dz**2))),
        ((1, nz), (1, nx), 0, 0, lambda j, i: 2 - (48 / 11) * A(j, i) * (1 / dx**2 + 1 / dz**2) + (30 / 11) * B(j, i) * (1 / dx**4 + 1 / dz**4 + 11 / (15 * dx**2 * dz**2))),
        ((1, nz), (1, nx - 1), 0, 1, lambda j, i: (24 / 11) * A(j, i) / dx**2 - B(j, i) * (15 / (11 * dx**4) + 1 / (dx**2 * dz**2))),
        ((1, nz - 1), (1, nx), 1, 0, lambda j, i: (24 / 11) * A(j, i) / dz**2 - B(j, i) * (15 / (11 * dz**4) + 1 / (dx**2 * dz**2))),
        ((2, nz), (2, nx), -1, -1, lambda j, i: B(j, i) / (2 * dx**2 * dz**2)),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: B(j, i) / (2 * dx**2 * dz**2)),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: B(j, i) / (2 * dx**2 * dz**2)),
        ((1, nz - 1), (1, nx - 1), 1, 1, lambda j, i: B(j, i) / (2 * dx**2 * dz**2)),
    ])
    # Auv_gen.m:24-78
    ops["Auv"] = _assemble(nz, nx, [
        ((1, nz), (2, nx), 0, -1, lambda j, i: (-20 / 33) * A(j, i) / dx + B(j, i) * (47 / (66 * dx**3) + 1 / (6 * dx * dz**2))),
        ((1, nz), (1, nx - 1), 0, 1, lambda j, i: (20 / 33) * A(j, i) / dx - B(j, i) * (47 / (66 * dx**3) + 1 / (6 * dx * dz**2))),
        ((2, nz), (2, nx), -1, -1, lambda j, i: -B(j, i) / (12 * dx * dz**2)),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: -B(j, i) / (12 * dx * dz**2)),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: B(j, i) / (12 * dx * dz**2)),
        ((1, nz - 1), (1, nx - 1), 1, 1, lambda j, i: B(j, i) / (12 * dx * dz**2)),
        ((1, nz), (3, nx), 0, -2, lambda j, i: (1 / 132) * A(j, i) / dx - (1 / 66) * B(j, i) / dx**3),
        ((1, nz), (1, nx - 2), 0, 2, lambda j, i: (-1 / 132) * A(j, i) / dx + (1 / 66) * B(j, i) / dx**3),
    ])
    # Auw_gen.m:24-78
    ops["Auw"] = _assemble(nz, nx, [
        ((1, nz - 1), (1, nx), 1, 0, lambda j, i: (20 / 33) * A(j, i) / dz - B(j, i) * (47 / (66 * dz**3) + 1 / (6 * dz * dx**2))),
        ((2, nz), (1, nx), -1, 0, lambda j, i: (-20 / 33) * A(j, i) / dz + B(j, i) * (47 / (66 * dz**3) + 1 / (6 * dz * dx**2))),
        ((2, nz), (2, nx), -1, -1, lambda j, i: -B(j, i) / (12 * dz * dx**2)),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: B(j, i) / (12 * dz * dx**2)),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: -B(j, i) / (12 * dz * dx**2)),
        ((1, nz - 1), (1, nx - 1), 1, 1, lambda j, i: B(j, i) / (12 * dz * dx**2)),
        ((1, nz - 2), (1, nx), 2, 0, lambda j, i: (-1 / 132) * A(j, i) / dz + (1 / 66) * B(j, i) / dz**3),
        ((3, nz), (1, nx), -2, 0, lambda j, i: (1 / 132) * A(j, i) / dz - (1 / 66) * B(j, i) / dz**3),
    ])

    # Avu_gen.m:25-65
    def P(j, i):
        return B(j, i) * (15 / (2 * dx**5) - 1 / (dx * dz**4) + 5 / (2 * dx**3 * dz**2))

    def Q(j, i):
        return A(j, i) * (15 / (2 * dx**3) - 1 / (dx * dz**2))

    def R(j, i):
        return B(j, i) * (5 / (4 * dx**3 * dz**2) - 1 / (2 * dx * dz**4)) + A(j, i) / (2 * dx * dz**2)

    ops["Avu"] = _assemble(nz, nx, [
        ((1, nz), (2, nx), 0, -1, lambda j, i: -P(j, i) + Q(j, i)),
        ((1, nz), (1, nx - 1), 0, 1, lambda j, i: P(j, i) - Q(j, i)),
        ((2, nz), (2, nx), -1, -1, R),
        ((1, nz - 1), (2, nx), 1, -1, R),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: -R(j, i)),
        ((1, nz - 1), (1, nx - 1), 1, 1, lambda j, i: -R(j, i)),
    ])
    # Avv_gen.m:24-85
    ops["Avv"] = _assemble(nz, nx, [
        ((2, nz), (1, nx), -1, 0, lambda j, i: A(j, i) / dz**2 - 2 * B(j, i) / (dx**2 * dz**2)),
        ((1, nz), (2, nx), 0, -1, lambda j, i: -(3 / 2) * A(j, i) / dx**2 + B(j, i) * (5 / (2 * dx**4) + 1 / (2 * dx**2 * dz**2))),
        ((1, nz), (1, nx), 0, 0, lambda j, i: 2 - A(j, i) * (12 / dx**2 + 2 / dz**2) + B(j, i) * (10 / dx**4 + 4 / (dx**2 * dz**2))),
        ((1, nz), (1, nx - 1), 0, 1, lambda j, i: -(3 / 2) * A(j, i) / dx**2 + B(j, i) * (5 / (2 * dx**4) + 1 / (2 * dx**2 * dz**2))),
        ((1, nz - 1), (1, nx), 1, 0, lambda j, i: A(j, i) / dz**2 - 2 * B(j, i) / (dx**2 * dz**2)),
        ((2, nz), (2, nx), -1, -1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
        ((1, nz - 1), (1, nx - 1), 1, 1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
    ])

    # Avw_gen.m:24-50
    def S(j, i):
        return B(j, i) / (4 * dz**3 * dx) - A(j, i) / (4 * dz * dx)

    ops["Avw"] = _assemble(nz, nx, [
        ((2, nz), (2, nx), -1, -1, S),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: -S(j, i)),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: -S(j, i)),
        ((1, nz - 1), (1, nx - 1), 1, 1, S),
    ])

    # Awu_gen.m:24-67
    def Pz(j, i):
        return B(j, i) * (15 / (2 * dz**5) - 1 / (dz * dx**4) + 5 / (2 * dz**3 * dx**2))

    def Qz(j, i):
        return A(j, i) * (15 / (2 * dz**3) - 1 / (dz * dx**2))

    def Rz(j, i):
        return B(j, i) * (5 / (4 * dz**3 * dx**2) - 1 / (2 * dz * dx**4)) + A(j, i) * (1 / (2 * dz * dx**2))

    ops["Awu"] = _assemble(nz, nx, [
        ((2, nz), (1, nx), -1, 0, lambda j, i: -Pz(j, i) + Qz(j, i)),
        ((1, nz - 1), (1, nx), 1, 0, lambda j, i: Pz(j, i) - Qz(j, i)),
        ((2, nz), (2, nx), -1, -1, Rz),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: -Rz(j, i)),
        ((2, nz), (1, nx - 1), -1, 1, Rz),
        ((1, nz - 1), (1, nx - 1), 1, 1, lambda j, i: -Rz(j, i)),
    ])

    # Awv_gen.m:26-52
    def Sx(j, i):
        return B(j, i) / (4 * dx**3 * dz) - A(j, i) / (4 * dz * dx)

    ops["Awv"] = _assemble(nz, nx, [
        ((2, nz), (2, nx), -1, -1, Sx),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: -Sx(j, i)),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: -Sx(j, i)),
        ((1, nz - 1), (1, nx - 1), 1, 1, Sx),
    ])
    # Aww_gen.m:26-87
    ops["Aww"] = _assemble(nz, nx, [
        ((2, nz), (1, nx), -1, 0, lambda j, i: -(3 / 2) * A(j, i) / dz**2 + B(j, i) * (5 / (2 * dz**4) + 1 / (2 * dx**2 * dz**2))),
        ((1, nz), (2, nx), 0, -1, lambda j, i: A(j, i) / dx**2 - 2 * B(j, i) / (dz**2 * dx**2)),
        ((1, nz), (1, nx), 0, 0, lambda j, i: 2 - A(j, i) * (12 / dz**2 + 2 / dx**2) + B(j, i) * (10 / dz**4 + 4 / (dx**2 * dz**2))),
        ((1, nz), (1, nx - 1), 0, 1, lambda j, i: A(j, i) / dx**2 - 2 * B(j, i) / (dz**2 * dx**2)),
        ((1, nz - 1), (1, nx), 1, 0, lambda j, i: -(3 / 2) * A(j, i) / dz**2 + B(j, i) * (5 / (2 * dz**4) + 1 / (2 * dx**2 * dz**2))),
        ((2, nz), (2, nx), -1, -1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
        ((1, nz - 1), (2, nx), 1, -1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
        ((2, nz), (1, nx - 1), -1, 1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
        ((1, nz - 1), (1, nx - 1), 1, 1, lambda j, i: -B(j, i) / (4 * dx**2 * dz**2)),
    ])
    return ops


# -------------------------------------------------------------------- CPML
def C_cof(R, d_pml, dx, dz, dt, nx, nz, v):
    """C_cof.m:4-65, full-grid arrays."""
    out = {}
    D = d_pml * dx
    x_b = np.concatenate([np.arange(d_pml, 0, -1) * dx, np.zeros(nx - 2 * d_pml),
                          np.arange(1, d_pml + 1) * dx])
    X = np.tile(x_b, (nz, 1))
    for side, mask in (("xl", np.s_[:, d_pml:]), ("xr", np.s_[:, :nx - d_pml])):
        Vs = v.copy()
        Vs[mask] = 0
        sig_ = ((3 * np.log(1 / R)) / (2 * D**3)) * Vs
        sig = sig_ * X**2
        dsig = 2 * sig_ * X
        den = 1 + sig * (dt / 2)
        out["c1" + side] = (1 - sig * (dt / 2)) / den
        out["c2" + side] = (dsig * dt) / den
        out["c3" + side] = (sig * dt) / den
    D = d_pml * dz
    z_b = np.concatenate([np.arange(d_pml, 0, -1) * dz, np.zeros(nz - 2 * d_pml),
                          np.arange(1, d_pml + 1) * dz])
    Z = np.tile(z_b[:, None], (1, nx))
    for side, mask in (("zt", np.s_[d_pml:, :]), ("zb", np.s_[:nz - d_pml, :])):
        Vs = v.copy()
        Vs[mask] = 0
        sig_ = ((3 * np.log(1 / R)) / (2 * D**3)) * Vs
        sig = sig_ * Z**2
        dsig = 2 * sig_ * Z
        den = 1 + sig * (dt / 2)
        out["c1" + side] = (1 - sig * (dt / 2)) / den
        out["c2" + side] = (dsig * dt) / den
        out["c3" + side] = (sig * dt) / den
    return out


def Ux123(M, U, dx, c):
    """Ux123_gen.m:4-17; M = dict of the six memories (updated in place)."""
    nz, nx = U.shape
    U_ax = np.zeros((nz, nx)); U_ax[:, :nx - 1] = U[:, 1:]
    U_dx = np.zeros((nz, nx)); U_dx[:, 1:] = U[:, :nx - 1]
    M["3l"] = c["c1xl"] * M["3l"] + (c["c2xl"] / (2 * dx)) * (U_dx - U_ax)
    M["2l"] = c["c1xl"] * M["2l"] + (c["c2xl"] / dx) * (U_dx - U_ax) + c["c3xl"] * ((U_dx - 2 * U + U_ax) / dx**2 - M["3l"])
    M["1l"] = c["c1xl"] * M["1l"] + (c["c2xl"] / (2 * dx)) * (U_dx - U_ax) + c["c3xl"] * (2 * (U_dx - 2 * U + U_ax) / dx**2 - M["2l"])
    M["3r"] = c["c1xr"] * M["3r"] + (c["c2xr"] / (2 * dx)) * (U_ax - U_dx)
    M["2r"] = c["c1xr"] * M["2r"] + (c["c2xr"] / dx) * (U_ax - U_dx) + c["c3xr"] * ((U_ax - 2 * U + U_dx) / dx**2 - M["3r"])
    M["1r"] = c["c1xr"] * M["1r"] + (c["c2xr"] / (2 * dx)) * (U_ax - U_dx) + c["c3xr"] * (2 * (U_ax - 2 * U + U_dx) / dx**2 - M["2r"])
    return M["1l"] + M["1r"]


def Uz123(M, U, dz, c):
    """Uz123_gen.m:4-17."""
    nz, nx = U.shape
    U_az = np.zeros((nz, nx)); U_az[:nz - 1, :] = U[1:, :]
    U_dz = np.zeros((nz, nx)); U_dz[1:, :] = U[:nz - 1, :]
    M["3t"] = c["c1zt"] * M["3t"] + (c["c2zt"] / (2 * dz)) * (U_dz - U_az)
    M["2t"] = c["c1zt"] * M["2t"] + (c["c2zt"] / dz) * (U_dz - U_az) + c["c3zt"] * ((U_dz - 2 * U + U_az) / dz**2 - M["3t"])
    M["1t"] = c["c1zt"] * M["1t"] + (c["c2zt"] / (2 * dz)) * (U_dz - U_az) + c["c3zt"] * (2 * (U_dz - 2 * U + U_az) / dz**2 - M["2t"])
    M["3b"] = c["c1zb"] * M["3b"] + (c["c2zb"] / (2 * dz)) * (U_az - U_dz)
    M["2b"] = c["c1zb"] * M["2b"] + (c["c2zb"] / dz) * (U_az - U_dz) + c["c3zb"] * ((U_az - 2 * U + U_dz) / dz**2 - M["3b"])
    M["1b"] = c["c1zb"] * M["1b"] + (c["c2zb"] / (2 * dz)) * (U_az - U_dz) + c["c3zb"] * (2 * (U_az - 2 * U + U_dz) / dz**2 - M["2b"])
    return M["1b"] + M["1t"]


def _mv(A, X):
    nz, nx = X.shape
    return (A @ X.ravel(order="F")).reshape(nz, nx, order="F")


def _mem(nz, nx, sides):
    return {k + s: np.zeros((nz, nx)) for k in "123" for s in sides}


def _update(ops, U, V, W, pU, pV, pW, V2dt2, F1):
    nU = _mv(ops["Auu"], U) + _mv(ops["Auv"], V) + _mv(ops["Auw"], W) - pU - V2dt2 * (F1["Ux"] + F1["Uz"])
    nV = _mv(ops["Avv"], V) + _mv(ops["Avu"], U) + _mv(ops["Avw"], W) - pV - V2dt2 * (F1["Vx"] + F1["Vz"])
    nW = _mv(ops["Aww"], W) + _mv(ops["Awu"], U) + _mv(ops["Awv"], V) - pW - V2dt2 * (F1["Wx"] + F1["Wz"])
    return nU, nV, nW


def fwd_PDE_exact_MNAD(dx, xs, xr, dz, zs, zr, v, dt, nt, w, d_pml, ops, indx=0):
    """fwd_PDE_exact_MNAD.m with 1-based xs, zs, xr, zr; returns
    (datau, U, V, W, UU, illum) -- UU/illum only when indx == 1."""
    V2dt2 = v**2 * (dt**2)
    nf = int(np.floor(nt / (2 * 3 * v.max() / v.min())))
    T = int(np.floor(nt / nf / 2)) if nf > 0 else None
    nz, nx = v.shape
    UU = np.zeros(((nz - 2 * d_pml) * (nx - 2 * d_pml), 2 * nf + 1))
    ttt = 1
    illum = 0.0
    datau = np.zeros((nt, len(xr)))
    pU, U, pV, V, pW, W = (np.zeros((nz, nx)) for _ in range(6))
    M = {f: (_mem(nz, nx, "lr"), _mem(nz, nx, "tb")) for f in "UVW"}
    c = C_cof(1e-4, d_pml, dx, dz, dt, nx, nz, v)
    for tt in range(1, nt + 1):
        F1 = {}
        for f, X in (("U", U), ("V", V), ("W", W)):
            F1[f + "x"] = Ux123(M[f][0], X, dx, c)
            F1[f + "z"] = Uz123(M[f][1], X, dz, c)
        U[zs - 1, xs - 1] = U[zs - 1, xs - 1] + V2dt2[zs - 1, xs - 1] * w[tt - 1]
        nU, nV, nW = _update(ops, U, V, W, pU, pV, pW, V2dt2, F1)
        for rr in range(len(xr)):
            datau[tt - 1, rr] = U[zr[rr] - 1, xr[rr] - 1]
        if indx == 1:
            if tt == 1:
                d2u = nU / dt**2
            elif tt == 2:
                d2u = (nU - 2 * U) / dt**2
            else:
                d2u = (nU - 2 * U + pU) / dt**2
            if tt % T == 0:
                internal = d2u[d_pml:nz - d_pml, d_pml:nx - d_pml]
                if ttt > UU.shape[1]:  # MATLAB grows UU on assignment
                    UU = np.concatenate([UU, np.zeros((UU.shape[0], ttt - UU.shape[1]))], axis=1)
                UU[:, ttt - 1] = internal.ravel(order="F")
                illum = illum + internal**2
                ttt += 1
        pU, U = U, nU
        pV, V = V, nV
        pW, W = W, nW
    if indx == 1:
        illum = T * illum
        return datau, U, V, W, UU, illum, T
    return datau, U, V, W, None, None, T


def Born_adj_MNAD(dx, dz, nt, dt, v, xr, zr, data, d_pml, UU, ops):
    """Born_adj_MNAD.m with U_r := U and data given as a matrix."""
    nz, nx = v.shape
    V2dt2 = (v * dt)**2
    nf = int(np.floor(nt / (2 * 3 * v.max() / v.min())))
    T = int(np.floor(nt / nf / 2))
    ttt = 1
    Im = 0.0
    datau = data
    pU, U, pV, V, pW, W = (np.zeros((nz, nx)) for _ in range(6))
    M = {f: (_mem(nz, nx, "lr"), _mem(nz, nx, "tb")) for f in "UVW"}
    c = C_cof(1e-4, d_pml, dx, dz, dt, nx, nz, v)
    for tt in range(nt, 0, -1):
        F1 = {}
        for f, X in (("U", U), ("V", V), ("W", W)):
            F1[f + "x"] = Ux123(M[f][0], X, dx, c)
            F1[f + "z"] = Uz123(M[f][1], X, dz, c)
        nU, nV, nW = _update(ops, U, V, W, pU, pV, pW, V2dt2, F1)
        for rr in range(len(xr)):
            U[zr[rr] - 1, xr[rr] - 1] = U[zr[rr] - 1, xr[rr] - 1] + V2dt2[zr[rr] - 1, xr[rr] - 1] * datau[tt - 1, rr]
        pU, U = U, nU
        pV, V = V, nV
        pW, W = W, nW
        if tt % T == 0:
            US = UU[:, UU.shape[1] - ttt]
            internal = U[d_pml:nz - d_pml, d_pml:nx - d_pml]
            Im = Im + US * internal.ravel(order="F")
            ttt += 1
    return T * Im


# ------------------------------------------------------------ driver steps
def ricker(fc, dt, nt):
    """main_Acoustic_Data.m:134-136 on t = (0:nt-1)*dt."""
    t = np.arange(nt) * dt
    t0 = 1.1 / fc
    tau = np.pi * (t - t0) * fc
    return (1 - 2 * tau**2) * np.exp(-tau**2)


def pad_edges(V, d_pml):
    """main_Acoustic_Data.m:144-145 (edge replication, x then z)."""
    v = np.concatenate([np.tile(V[:, :1], (1, d_pml)), V, np.tile(V[:, -1:], (1, d_pml))], axis=1)
    return np.concatenate([np.tile(v[:1, :], (d_pml, 1)), v, np.tile(v[-1:, :], (d_pml, 1))], axis=0)


def v_smooth(v, s1, s2, nit):
    """v_smooth.m: nit passes of an s1 x s2 box filter on slowness."""
    s0 = 1.0 / v
    hh1, hh2 = (s1 - 1) // 2, (s2 - 1) // 2
    n1, n2 = s0.shape
    for _ in range(nit):
        p = np.concatenate([np.tile(s0[:, :1], (1, hh2)), s0, np.tile(s0[:, -1:], (1, hh2))], axis=1)
        p = np.concatenate([np.tile(p[:1, :], (hh1, 1)), p, np.tile(p[-1:, :], (hh1, 1))], axis=0)
        s = np.zeros((n1, n2))
        for i2 in range(n2):
            for i1 in range(n1):
                s[i1, i2] = p[i1:i1 + s1, i2:i2 + s2].sum() / (s1 * s2)
        s0 = s
    return 1.0 / s0


# ---------------------------------------------------------------- fixtures
def _case_forward(name, V, d_pml, dx, dz, dt, nt, fc, src, recs):
    v = pad_edges(V, d_pml)
    ops = build_ops(dx, dz, dt, v)
    w = ricker(fc, dt, nt)
    zs, xs = src
    zr = np.array([r[0] for r in recs]); xr = np.array([r[1] for r in recs])
    datau, U, V_, W, _, _, _ = fwd_PDE_exact_MNAD(dx, xs, xr, dz, zs, zr, v, dt, nt, w, d_pml, ops)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), v=v, w=w, dx=dx, dz=dz, dt=dt,
                        npml=d_pml, zs=zs - 1, xs=xs - 1, zr=zr - 1, xr=xr - 1,
                        gather=datau, U=U, V=V_, W=W)
    print(name, "gather max", np.abs(datau).max())


def _case_rtm(name, V, d_pml, dx, dz, dt, nt, fc, shots):
    """Two-model RTM of main_Acoustic_RTM.m:172-262 on a tiny grid."""
    v = pad_edges(V, d_pml)
    Vs = v_smooth(V, 3, 3, 3)
    vs = pad_edges(Vs, d_pml)
    ops_t = build_ops(dx, dz, dt, v)
    ops_s = build_ops(dx, dz, dt, vs)
    w = ricker(fc, dt, nt)
    nzi, nxi = V.shape
    out = dict(v=v, vs=vs, Vs=Vs, w=w, dx=dx, dz=dz, dt=dt, npml=d_pml, nshots=len(shots))
    ims, ills = [], []
    for k, (src, recs) in enumerate(shots):
        zs, xs = src
        zr = np.array([r[0] for r in recs]); xr = np.array([r[1] for r in recs])
        data = fwd_PDE_exact_MNAD(dx, xs, xr, dz, zs, zr, v, dt, nt, w, d_pml, ops_t)[0]
        datas, _, _, _, UU, illum, T = fwd_PDE_exact_MNAD(dx, xs, xr, dz, zs, zr, vs, dt, nt, w, d_pml, ops_s, 1)
        sct = data - datas
        Im = Born_adj_MNAD(dx, dz, nt, dt, vs, xr, zr, sct, d_pml, UU, ops_s)
        out.update({f"zs{k}": zs - 1, f"xs{k}": xs - 1, f"zr{k}": zr - 1, f"xr{k}": xr - 1,
                    f"data{k}": data, f"datas{k}": datas, f"UU{k}": UU, f"illum{k}": illum,
                    f"Im{k}": Im})
        ims.append(Im.reshape(nzi, nxi, order="F"))
        ills.append(illum)
        out["T"] = T
        out["ncol"] = UU.shape[1]
    # main_Acoustic_RTM.m:237-244,256-262
    illum_mean = sum(ills) / len(ills)
    img = np.mean(np.stack(ims, axis=2), axis=2) / illum_mean
    out["img"] = img
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "T", out["T"], "ncol", out["ncol"], "img max", np.abs(img).max())


def main():
    # Forward, dx == dz, two layers, source next to the top strip (the
    # physical row 1 of main_Acoustic_Data.m:120-126), 5 receivers on row 2.
    V = np.full((23, 31), 1500.0)
    V[12:, :] = 2500.0
    d = 6
    _case_forward("fwd_square", V, d, 10.0, 10.0, 1e-3, 160, 25.0,
                  (d + 1, d + 16), [(d + 2, d + c) for c in (1, 8, 16, 24, 31)])
    # Forward, dx != dz, source inside the top-left corner strip, receivers
    # spread over rows and strips (duplicated receiver included).
    V = np.full((19, 27), 1800.0)
    V[:, 14:] = 2200.0
    V[10:, :] += 400.0
    d = 5
    _case_forward("fwd_rect", V, d, 12.0, 9.0, 1.2e-3, 140, 20.0,
                  (3, 4), [(2, 2), (d + 3, d + 5), (d + 3, d + 5), (d + 10, d + 20), (24, 30)])
    # RTM, two shots; reflector at row 9.  nt chosen so that ncol = K + 1.
    V = np.full((16, 24), 1500.0)
    V[9:, :] = 2000.0
    d = 5
    shots = []
    for xs in (7, 17):
        shots.append(((d + 1, d + xs), [(d + 2, d + c) for c in range(1, 25)]))
    _case_rtm("rtm_two_shot", V, d, 10.0, 10.0, 1e-3, 200, 30.0, shots)
    # RTM, one shot, dx != dz, nt chosen so that ncol == K.
    V = np.full((14, 20), 1600.0)
    V[7:, :] = 2400.0
    d = 4
    _case_rtm("rtm_rect", V, d, 11.0, 9.0, 1e-3, 187, 30.0,
              [((d + 2, d + 10), [(d + 1, d + c) for c in range(1, 21)])])


if __name__ == "__main__":
    main()

"""ctypes wrapper of the FP64 oracle (oracle/mnad_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product library never does.

All grids are numpy float64 arrays, row-major [nz][nx] (x fastest), with
0-based indices; see mnad_oracle.h for the reference lines each entry restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmnad_oracle.so")
_lib = None


class _Grid(C.Structure):
    _fields_ = [("nz", C.c_int), ("nx", C.c_int), ("npml", C.c_int),
                ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc) and return the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        gp = C.POINTER(_Grid)
        L.mo_snap_schedule.argtypes = [C.c_int, C.c_double, C.c_double, ip, ip, ip, ip]
        L.mo_cpml_coeffs.argtypes = [gp, dp, C.c_double, C.c_int, dp, dp, dp]
        L.mo_apply_A.argtypes = [gp, dp, dp, dp, dp, dp, dp, dp]
        L.mo_forward.argtypes = [gp, dp, C.c_int, dp, C.c_int, C.c_int, C.c_int, ip, ip,
                                 C.c_int, dp, dp, dp, dp, dp, dp]
        L.mo_adjoint.argtypes = [gp, dp, C.c_int, C.c_int, ip, ip, dp, dp, dp]
        L.mo_steps.argtypes = [gp, dp, C.c_int, dp, dp, dp, dp, dp, dp]
        for f in (L.mo_snap_schedule, L.mo_forward, L.mo_adjoint, L.mo_steps):
            f.restype = C.c_int
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def grid(nz, nx, npml, dx, dz, dt):
    return _Grid(int(nz), int(nx), int(npml), float(dx), float(dz), float(dt))


def snap_schedule(nt, vmin, vmax):
    """(nf, T, K, ncol) of fwd_PDE_exact_MNAD.m:46-47,53; None if nf == 0."""
    a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = lib().mo_snap_schedule(int(nt), float(vmin), float(vmax),
                                C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    if rc:
        return None
    return a.value, b.value, c.value, d.value


def cpml_coeffs(g, v, side, R=1e-4):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = [np.empty_like(v) for _ in range(3)]
    lib().mo_cpml_coeffs(C.byref(g), _dp(v), float(R), int(side), *[_dp(o) for o in out])
    return out


def apply_A(g, v, U, V, W):
    v = np.ascontiguousarray(v, dtype=np.float64)
    ins = [np.ascontiguousarray(x, dtype=np.float64) for x in (U, V, W)]
    outs = [np.empty_like(v) for _ in range(3)]
    lib().mo_apply_A(C.byref(g), _dp(v), *[_dp(x) for x in ins], *[_dp(o) for o in outs])
    return outs


def forward(g, v, w, zs, xs, zr, xr, snap=False, want_state=True):
    """Forward shot; returns dict(gather, U, V, W[, UU, illum, T, K, ncol])."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64).ravel()
    nt = w.size
    zr = np.ascontiguousarray(zr, dtype=np.int32).ravel()
    xr = np.ascontiguousarray(xr, dtype=np.int32).ravel()
    nr = zr.size
    gather = np.zeros((nt, nr))
    U = V = W = None
    if want_state:
        U, V, W = (np.zeros_like(v) for _ in range(3))
    res = dict(gather=gather, U=U, V=V, W=W)
    UU = illum = None
    if snap:
        sch = snap_schedule(nt, v.min(), v.max())
        if sch is None:
            raise ValueError("nt too short for the snapshot schedule (nf == 0)")
        nf, T, K, ncol = sch
        ni = (g.nz - 2 * g.npml) * (g.nx - 2 * g.npml)
        UU = np.zeros((K, ni))
        illum = np.zeros(ni)
        res.update(UU=UU, illum=illum, T=T, K=K, ncol=ncol, nf=nf)
    rc = lib().mo_forward(C.byref(g), _dp(v), nt, _dp(w), int(zs), int(xs), nr,
                          _ip(zr), _ip(xr), int(bool(snap)), _dp(gather),
                          _dp(U), _dp(V), _dp(W), _dp(UU), _dp(illum))
    if rc:
        raise ValueError(f"mo_forward failed ({rc})")
    return res


def adjoint(g, v, zr, xr, data, UU):
    """RTM adjoint with imaging; returns Im [nzi*nxi] (row-major interior)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    nt = data.shape[0]
    zr = np.ascontiguousarray(zr, dtype=np.int32).ravel()
    xr = np.ascontiguousarray(xr, dtype=np.int32).ravel()
    UU = np.ascontiguousarray(UU, dtype=np.float64)
    ni = (g.nz - 2 * g.npml) * (g.nx - 2 * g.npml)
    Im = np.zeros(ni)
    rc = lib().mo_adjoint(C.byref(g), _dp(v), nt, zr.size, _ip(zr), _ip(xr),
                          _dp(data), _dp(UU), _dp(Im))
    if rc:
        raise ValueError(f"mo_adjoint failed ({rc})")
    return Im


def steps(g, v, nsteps, prev, cur):
    """Bare updates (CPML on, no source/receivers) from (prev, cur) states."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    p = [np.array(x, dtype=np.float64, order="C") for x in prev]
    c = [np.array(x, dtype=np.float64, order="C") for x in cur]
    rc = lib().mo_steps(C.byref(g), _dp(v), int(nsteps), *[_dp(x) for x in p],
                        *[_dp(x) for x in c])
    if rc:
        raise ValueError(f"mo_steps failed ({rc})")
    return p, c

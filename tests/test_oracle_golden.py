"""The FP64 oracle (stencil form, oracle/mnad_oracle.c) against the golden
fixtures of the independent sparse-matrix transliteration
(tests/golden/make_golden.py).  Both restate the same MATLAB files by
different routes: agreement to ~1e-13 pins indices, signs and step order."""
import numpy as np
import pytest

from conftest import golden

TOL = 1e-12


def _grid(O, f, vkey="v"):
    v = f[vkey]
    return O.grid(v.shape[0], v.shape[1], int(f["npml"]), float(f["dx"]), float(f["dz"]),
                  float(f["dt"]))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", ["fwd_square", "fwd_rect"])
def test_forward_matches_sparse_transliteration(oracle, name):
    f = golden(name)
    g = _grid(oracle, f)
    r = oracle.forward(g, f["v"], f["w"], int(f["zs"]), int(f["xs"]), f["zr"], f["xr"])
    for k in ("gather", "U", "V", "W"):
        assert _rel(r[k], f[k]) < TOL, k


@pytest.mark.parametrize("name", ["rtm_two_shot", "rtm_rect"])
def test_rtm_matches_sparse_transliteration(oracle, name):
    """Snapshots, illumination, pairing shift (ncol - K) and the adjoint."""
    f = golden(name)
    g = _grid(oracle, f, "vs")
    nzi, nxi = g.nz - 2 * g.npml, g.nx - 2 * g.npml
    ims, ills = [], []
    for k in range(int(f["nshots"])):
        zs, xs, zr, xr = (int(f[f"zs{k}"]), int(f[f"xs{k}"]), f[f"zr{k}"], f[f"xr{k}"])
        obs = oracle.forward(g, f["v"], f["w"], zs, xs, zr, xr, want_state=False)
        assert _rel(obs["gather"], f[f"data{k}"]) < TOL
        syn = oracle.forward(g, f["vs"], f["w"], zs, xs, zr, xr, snap=True, want_state=False)
        assert syn["T"] == int(f["T"]) and syn["ncol"] == int(f["ncol"])
        assert _rel(syn["gather"], f[f"datas{k}"]) < TOL
        if f"UU{k}" in f:
            UUg = f[f"UU{k}"]                       # [ni x ncol], MATLAB column-major interior
            UUo = syn["UU"].reshape(syn["K"], nzi, nxi).transpose(0, 2, 1).reshape(syn["K"], -1)
            assert _rel(UUo, UUg[:, :syn["K"]].T) < TOL
            assert np.all(UUg[:, syn["K"]:] == 0)   # columns never written stay zero
        ill = syn["illum"].reshape(nzi, nxi)
        assert _rel(ill, f[f"illum{k}"]) < TOL
        Im = oracle.adjoint(g, f["vs"], zr, xr, obs["gather"] - syn["gather"], syn["UU"])
        Img = f[f"Im{k}"].reshape(nxi, nzi).T
        assert _rel(Im.reshape(nzi, nxi), Img) < TOL
        ims.append(Im.reshape(nzi, nxi))
        ills.append(ill)
    img = np.mean(ims, axis=0) / (sum(ills) / len(ills))     # main_Acoustic_RTM.m:243,262
    assert _rel(img, f["img"]) < 1e-11


def test_snapshot_schedule_closed_forms(oracle):
    """fwd_PDE_exact_MNAD.m:46-47,53 on the schedule values the configs and
    the shipped example imply (C1, C3, C5, examples/*: 1500-4440 m/s, nt 1922)."""
    assert oracle.snap_schedule(1000, 1500, 2500) == (100, 5, 200, 201)
    assert oracle.snap_schedule(6000, 1500, 4500) == (333, 9, 666, 667)
    assert oracle.snap_schedule(12000, 1500, 4500) == (666, 9, 1333, 1333)
    assert oracle.snap_schedule(1922, 1500, 4440) == (108, 8, 240, 240)
    assert oracle.snap_schedule(5, 1500, 1500) is None     # nf = 0 -> T = Inf
    # T >= 3 always (nt/nf >= 6 vmax/vmin >= 6), so the tt = 1, 2 special
    # cases of d2u (fwd_PDE_exact_MNAD.m:128-131) never reach a snapshot.
    for nt in range(12, 400, 7):
        for r in (1.0, 1.3, 2.0, 3.7):
            s = oracle.snap_schedule(nt, 1000.0, 1000.0 * r)
            if s is not None:
                assert s[1] >= 3

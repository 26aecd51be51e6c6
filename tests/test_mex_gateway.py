"""The MEX gateway compiled against the mock mx API (no MATLAB here):
command dispatch, argument validation with the reference's 1-based indices,
and the mnad:<kind> error identifiers of the C-ABI status codes."""
import numpy as np
import pytest

from conftest import golden


@pytest.fixture(scope="module")
def mex(mnad):
    from mexgw import Mex
    return Mex()


def _fwd_args(f, **over):
    a = dict(dx=float(f["dx"]), dz=float(f["dz"]), dt=float(f["dt"]), d_pml=int(f["npml"]),
             v=f["v"], w=f["w"], nt=f["w"].size, xs=int(f["xs"]) + 1, zs=int(f["zs"]) + 1,
             xr=f["xr"] + 1, zr=f["zr"] + 1, ss=1, indx=0)
    a.update(over)
    return ["fwd", a["dx"], a["dz"], a["dt"], a["d_pml"], a["v"], a["w"], a["nt"], a["xs"],
            a["zs"], a["xr"], a["zr"], a["ss"], a["indx"]]


def test_version(mex):
    assert "gfx950" in mex("version")


def test_unknown_command_and_arity(mex):
    from mexgw import MexError
    with pytest.raises(MexError) as e:
        mex("bogus")
    assert e.value.ident == "mnad:arg"
    with pytest.raises(MexError) as e:
        mex("fwd", 1.0, 2.0)
    assert e.value.ident == "mnad:arg"


@pytest.mark.parametrize("bad", [dict(xs=3.5), dict(zs=0), dict(nt=7), dict(xr=[1.5, 2])])
def test_fwd_argument_validation(mex, bad):
    """Non-integer or non-positive 1-based indices and length(w) != nt are
    rejected before any device work."""
    from mexgw import MexError
    f = golden("fwd_square")
    with pytest.raises(MexError) as e:
        mex(*_fwd_args(f, **bad))
    assert e.value.ident == "mnad:arg"


def test_adj_data_shape_validation(mex):
    from mexgw import MexError
    f = golden("rtm_rect")
    with pytest.raises(MexError) as e:
        mex("adj", float(f["dx"]), float(f["dz"]), f["w"].size, float(f["dt"]), f["vs"],
            f["xr0"] + 1, f["zr0"] + 1, np.zeros((3, 3)), int(f["npml"]), 1)
    assert e.value.ident == "mnad:arg"


def test_valid_call_without_gpu_raises_nogpu(mex, mnad):
    from mexgw import MexError
    if mnad.lib().mnad_device_count() > 0:
        pytest.skip("a GPU is visible here")
    f = golden("fwd_square")
    with pytest.raises(MexError) as e:
        mex(*_fwd_args(f))
    assert e.value.ident == "mnad:nogpu"


def test_batched_and_serving_argument_validation(mex):
    """The batched and serving commands check their arguments before any
    device work: receiver counts that do not add up, v and vs of different
    sizes, a wrong arity."""
    from mexgw import MexError
    f = golden("rtm_two_shot")
    dx, dz, dt, d = float(f["dx"]), float(f["dz"]), float(f["dt"]), int(f["npml"])
    w = f["w"]
    xs, zs = [int(f["xs0"]) + 1, int(f["xs1"]) + 1], [int(f["zs0"]) + 1, int(f["zs1"]) + 1]
    xr = np.concatenate([f["xr0"], f["xr1"]]) + 1
    zr = np.concatenate([f["zr0"], f["zr1"]]) + 1
    with pytest.raises(MexError) as e:   # sum(nrec) != length(xr)
        mex("fwd_shots", dx, dz, dt, d, f["v"], w, w.size, xs, zs, xr, zr, [3, 3], 0)
    assert e.value.ident == "mnad:arg"
    with pytest.raises(MexError) as e:   # vs of another size
        mex("rtm_shots", dx, dz, dt, d, f["v"], f["vs"][:-1], w, w.size, xs, zs, xr, zr,
            [f["xr0"].size, f["xr1"].size], 0)
    assert e.value.ident == "mnad:arg"
    with pytest.raises(MexError) as e:
        mex("submit", 0, dx, dz, dt)
    assert e.value.ident == "mnad:arg"
    with pytest.raises(MexError) as e:
        mex("wait", 1)
    assert e.value.ident == "mnad:arg"

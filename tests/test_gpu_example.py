"""The reference's own example workload on the GPU: examples/modeling/
main_Acoustic_Data.m on its V_Hess model (the reference's only data
fixture, decimated and scaled as the driver does: 200 x 500, 1500-4440 m/s),
fc 15 Hz, dx = 13 m, dt = 1.5616 ms, nt = 1922 (t = 0:dt:3), d_pml 40, the
fixspread geometry of acquisi.m (one source at 1-based padded (41, 290),
499 receivers on row 42), through the MATLAB-signature mirror
fwd_PDE_exact_MNAD with the driver's 1-based indices.  The gather must meet
the FP64 oracle run on the same inputs at 1e-4."""
import sys

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu


def test_example_modelling_workload_vs_oracle(mnad, oracle):
    import importlib
    synth = importlib.import_module("mnad_amd.synth")
    V = golden("v_hess_example")["V"]
    nz, nx = V.shape
    d_pml = 40
    dx, dt, nt = synth.reference_dx_dt(V.min(), V.max(), 15.0, 3.0, with_end=True)
    src, rec = synth.acquisi(V, dx, dx, 1 * dx, 2 * dx, nx, (nx / 2 + 1) * dx, dx, (nx / 2) * dx, dx)
    xs, zs, xr, zr = synth.to_padded_indices(src, rec, dx, dx, d_pml)
    v = synth.pad_edges(V, d_pml)
    w = synth.ricker(15.0, dt, nt)
    print(f"example workload {v.shape[0]}x{v.shape[1]} nt={nt} dx={dx} dt={dt:.6g}", flush=True)
    sys.stdout.flush()
    data = mnad.fwd_PDE_exact_MNAD(dx, xs[0], xr[0], dx, zs[0], zr[0], v, dt, nt, w, d_pml, 1,
                                   *([None] * 9))
    g = oracle.grid(v.shape[0], v.shape[1], d_pml, dx, dx, dt)
    ref = oracle.forward(g, v, w, int(zs[0]) - 1, int(xs[0]) - 1,
                         np.asarray(zr[0], int) - 1, np.asarray(xr[0], int) - 1, want_state=False)
    err = rel_l2(data[0], ref["gather"])
    print(f"  gather rel L2 vs FP64 {err:.2e}", flush=True)
    assert data[0].shape == (1922, 499)
    assert err < 1e-4

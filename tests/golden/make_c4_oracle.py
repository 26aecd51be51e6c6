"""Writes tests/golden/c4_oracle.npz: FP64 oracle results of two requests of
the seeded C4 serving slice the GPU suite serves (bench.C4Requests(12, 41):
grids U[256,2048] x U[256,1024] + 20 CPML cells, nt U[1000,8000], 70 %
modelling / 30 % RTM).  Run in the build container (a few minutes of CPU):

    python tests/golden/make_c4_oracle.py

Stored, as float32 (mantissas cut to 18 bits, relative rounding <= 2e-6):
  - request 0 (modelling, the smallest): every 8th receiver of its gather
    (fwd_PDE_exact_MNAD.m:119-123);
  - request 6 (RTM): its per-shot image Im (Born_adj_MNAD.m:138) and
    illumination on every 3rd row and column of the interior, from the
    observed-data forward, the migration-model forward with d2u snapshots and
    the adjoint (main_Acoustic_RTM.m:223-256).
SHA-256 digests of the regenerated models and wavelets are stored beside."""
from __future__ import annotations

import os
import sys
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_full_size import cut, digest  # noqa: E402

SEED, NREQ, MODEL_REQ, RTM_REQ, RSTEP, ISTEP = 41, 12, 0, 6, 8, 3


def _forward(args):
    k, which = args
    import bench
    from oracle import oracle as O
    nz, nx, npml, dx, dt, v, vs, w, zs, xs, zr, xr = bench.C4Requests(NREQ, SEED).build(k)
    g = O.grid(nz, nx, npml, dx, dx, dt)
    m = v if which == "true" else vs
    return O.forward(g, m, w, zs, xs, zr, xr, snap=(which == "smooth"), want_state=False)


def main():
    import bench
    from oracle import oracle as O
    reqs = bench.C4Requests(NREQ, SEED)
    assert not reqs.specs[MODEL_REQ]["rtm"] and reqs.specs[RTM_REQ]["rtm"]
    with Pool(3) as pool:
        mod, obs, syn = pool.map(_forward, [(MODEL_REQ, "true"), (RTM_REQ, "true"),
                                            (RTM_REQ, "smooth")])
    nz, nx, npml, dx, dt, v, vs, w, zs, xs, zr, xr = reqs.build(RTM_REQ)
    g = O.grid(nz, nx, npml, dx, dx, dt)
    Im = O.adjoint(g, vs, zr, xr, obs["gather"] - syn["gather"], syn["UU"])
    nzi, nxi = nz - 2 * npml, nx - 2 * npml
    Im = Im.reshape(nzi, nxi)
    ill = syn["illum"].reshape(nzi, nxi)
    m = reqs.build(MODEL_REQ)
    out = dict(seed=SEED, nreq=NREQ, model_req=MODEL_REQ, rtm_req=RTM_REQ, rstep=RSTEP,
               istep=ISTEP, gather=cut(mod["gather"][:, ::RSTEP], 18),
               Im=cut(Im[::ISTEP, ::ISTEP], 18), illum=cut(ill[::ISTEP, ::ISTEP], 18),
               sha_model_v=digest(m[5]), sha_model_w=digest(m[7]), sha_rtm_v=digest(v),
               sha_rtm_vs=digest(vs), sha_rtm_w=digest(w))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_oracle.npz")
    np.savez_compressed(path, **out)
    print(f"# wrote {path} ({os.path.getsize(path) / 2**20:.2f} MiB)", flush=True)


if __name__ == "__main__":
    main()

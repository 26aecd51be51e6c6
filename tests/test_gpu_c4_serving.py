"""The C4 serving workload at its real sizes on the GPU: a seeded slice of 12
requests drawn from C4's distribution (bench.C4Requests(12, 41): physical
grids U[256,2048] x U[256,1024] + 20 CPML cells, nt U[1000,8000], Ricker
U[10,30] Hz, 70 % modelling / 30 % RTM) through the continuous-batching
server with the bench's admission settings.  Each request is served alone,
then all twelve are submitted at once and overlap on the device; every
result in the mix must equal its result alone bitwise (a shot's arithmetic
does not depend on its neighbours in a launch).  The smallest modelling
request (gather, 1e-4) and an RTM request (per-shot image Im, f16 snapshots,
1e-3) are checked against FP64 oracle results computed in the build
container (tests/golden/make_c4_oracle.py)."""
import hashlib
import sys

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes(order="F")).hexdigest()


def test_c4_slice_alone_equals_mix_and_meets_the_oracle(mnad):
    import bench
    f = golden("c4_oracle")
    reqs = bench.C4Requests(int(f["nreq"]), int(f["seed"]))
    n = len(reqs.specs)
    built = [reqs.build(k) for k in range(n)]
    m, r = built[int(f["model_req"])], built[int(f["rtm_req"])]
    assert _digest(m[5]) == str(f["sha_model_v"]) and _digest(m[7]) == str(f["sha_model_w"])
    assert _digest(r[5]) == str(f["sha_rtm_v"]) and _digest(r[6]) == str(f["sha_rtm_vs"])
    assert _digest(r[7]) == str(f["sha_rtm_w"])
    ctx = mnad.Context(0)
    try:
        ctx.serve_start(hbm_budget=0, max_active=64, chunk_steps=32, max_points=int(4e6))

        def submit(k):
            nz, nx, npml, dx, dt, v, vs, w, zs, xs, zr, xr = built[k]
            kind = mnad.REQ_RTM if vs is not None else mnad.REQ_MODEL
            return ctx.submit(kind, mnad.Grid(nz, nx, npml, dx, dx, dt), v, w, zs, xs, zr, xr,
                              v_smooth=vs, snap_fmt=mnad.SNAP_F16, want_gather=vs is None)

        alone = []
        for k in range(n):
            out, info = ctx.wait(submit(k))
            alone.append(out)
            sp = reqs.specs[k]
            _say(f"  request {k} alone: {sp['nzp']}x{sp['nxp']} nt {sp['nt']} "
                 f"{'RTM' if sp['rtm'] else 'modelling'}, {info['done_s'] - info['admit_s']:.3f} s")
        tickets = [submit(k) for k in range(n)]
        mix = [ctx.wait(t) for t in tickets]
        ctx.serve_stop()
    finally:
        ctx.close()
    spans = sorted((info["admit_s"], info["done_s"]) for _, info in mix)
    overlaps = sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1])
    _say(f"  mix: {overlaps} of {n - 1} consecutive admissions overlap a running request")
    assert overlaps >= 1
    for k in range(n):
        key = "image" if reqs.specs[k]["rtm"] else "gather"
        assert np.array_equal(mix[k][0][key], alone[k][key]), k
    mk, rk = int(f["model_req"]), int(f["rtm_req"])
    e_g = rel_l2(alone[mk]["gather"][:, ::int(f["rstep"])], f["gather"])
    s = int(f["istep"])
    e_i = rel_l2(alone[rk]["image"][::s, ::s], f["Im"])
    _say(f"  vs FP64: request {mk} gather {e_g:.2e}, request {rk} image {e_i:.2e}")
    assert e_g < 1e-4
    assert e_i < 1e-3

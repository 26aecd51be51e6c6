"""Full-size parity check of the C2 workload (561 x 1061 padded, nt = 4000),
or of the 1001 x 501 window of the C5 model (--workload c5crop, nt = 12000),
against the FP64 oracle.  Not collected by pytest (minutes of CPU): run it
on the outputs `bench.py --save-outputs FILE` wrote on the GPU:

    python tests/validate_full_size.py FILE [MORE_FILES...] [--nt 4000]

It regenerates the same seeded inputs, runs the oracle's observed-data
forward, migration-model forward with snapshots, and adjoint (the first two
in parallel processes), and prints relative L2 errors of the gathers, the
scattered data, the image img = Im./illum and the illumination as JSON."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _workload(name, nt):
    import bench
    return bench.c5_workload(nt, 1, crop=True) if name == "c5crop" else bench.c2_workload(nt)


def _forward(args):
    which, nt, snap, name = args
    from oracle import oracle as O
    wl = _workload(name, nt)
    v = wl["v"] if which == "true" else wl["vs"]
    g = O.grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    t0 = time.time()
    r = O.forward(g, v, wl["w"], wl["zs"], wl["xs"], wl["zr"], wl["xr"], snap=snap,
                  want_state=False)
    print(f"# oracle forward ({which}) {time.time() - t0:.0f} s", file=sys.stderr, flush=True)
    return r


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("outputs")
    ap.add_argument("more", nargs="*", help="further output files (other snapshot formats)")
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--workload", default="c2", choices=["c2", "c5crop"])
    args = ap.parse_args()
    from oracle import oracle as O
    with Pool(2) as pool:
        obs, syn = pool.map(_forward, [("true", args.nt, False, args.workload),
                                       ("smooth", args.nt, True, args.workload)])
    wl = _workload(args.workload, args.nt)
    vs = wl["vs"]
    g = O.grid(vs.shape[0], vs.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    t0 = time.time()
    Im = O.adjoint(g, vs, wl["zr"], wl["xr"], obs["gather"] - syn["gather"], syn["UU"])
    print(f"# oracle adjoint {time.time() - t0:.0f} s", file=sys.stderr, flush=True)
    nzi, nxi = g.nz - 2 * g.npml, g.nx - 2 * g.npml
    ill = syn["illum"].reshape(nzi, nxi)
    img = Im.reshape(nzi, nxi) / ill
    for path in [args.outputs] + args.more:
        gpu = np.load(path)
        res = {"workload": "%s %dx%d nt=%d snapshots=%s" % (args.workload, g.nz, g.nx, args.nt,
                                                            str(gpu["snap"]))}
        if "dobs" in gpu:
            res.update(
                gather_obs_rel_l2=rel(gpu["dobs"], obs["gather"]),
                gather_syn_rel_l2=rel(gpu["dsyn"], syn["gather"]),
                scattered_rel_l2=rel(gpu["dobs"].astype(float) - gpu["dsyn"],
                                     obs["gather"] - syn["gather"]))
        res.update(illum_rel_l2=rel(gpu["illum"], ill), image_rel_l2=rel(gpu["img"], img),
                   image_Im_rel_l2=rel(gpu["img"] * gpu["illum"], Im.reshape(nzi, nxi)))
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

"""Step-kernel time per step (HIP events on the library stream, from
mnad_timing) of several kernel variants on forward batches of a given grid.

    python3 mnad-modeling-rtm_amd/microbench/step_kernel_compare.py \
        --grid 561x1061x30 --shots 1,2 --kernels 4:1:1,9:2:8 --nt 1000

Each variant runs one untimed call (set-up, graph capture) and then
--calls timed calls of fwd_shots on a linear-gradient model (1500 to
3000 m/s), source at the centre of the first physical row, a receiver on
every interior column of the next row.  --calls 1 --warm 0 makes one call
per variant (for profiler runs)."""
import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", default="561x1061x30", help="padded nz x nx x npml, comma list")
    ap.add_argument("--shots", default="1,2", help="shots per launch, comma list")
    ap.add_argument("--kernels", default="4:1:1,9:2:8",
                    help="kernel:rows_per_thread:block_rows, comma list")
    ap.add_argument("--nt", type=int, default=1000)
    ap.add_argument("--calls", type=int, default=1)
    ap.add_argument("--warm", type=int, default=1)
    a = ap.parse_args()
    from __graft_entry__ import load_package
    mnad = load_package()
    dt, fc = 8e-4, 20.0
    t = np.arange(a.nt) * dt
    tau = np.pi * (t - 1.1 / fc) * fc
    w = (1 - 2 * tau**2) * np.exp(-tau**2)
    ctx = mnad.Context(0)
    try:
        for gs in a.grid.split(","):
            nz, nx, npml = (int(x) for x in gs.split("x"))
            v = np.asfortranarray(1500.0 + 1500.0 * np.arange(nz)[:, None] / nz * np.ones((1, nx)))
            g = mnad.Grid(nz, nx, npml, 10.0, 10.0, dt)
            xr = np.arange(npml, nx - npml)
            zr = np.full(xr.size, npml + 1)
            for ks in a.kernels.split(","):
                k, r, b = (int(x) for x in ks.split(":"))
                ctx.set_tuning(kernel=k, rows_per_thread=r, block_rows=b)
                for ns in (int(x) for x in a.shots.split(",")):
                    xs = [nx // 2 + 3 * q for q in range(ns)]
                    args = (g, v, w, [npml] * ns, xs, [(zr, xr)] * ns)
                    try:
                        for _ in range(a.warm):
                            ctx.fwd_shots(*args, batch=ns)
                        ctx.timing(reset=True)
                        for _ in range(a.calls):
                            ctx.fwd_shots(*args, batch=ns)
                        ms, launches, _ = ctx.timing(reset=True)
                        us = ms * 1e3 / (a.nt * a.calls)
                        print(f"{nz}x{nx} kernel {k},{r},{b} {ns} shot(s): {us:.2f} us per step, "
                              f"{us / ns:.2f} us per shot-step ({launches} launches)", flush=True)
                    except mnad.MnadError as e:
                        print(f"{nz}x{nx} kernel {k},{r},{b} {ns} shot(s): {e}", flush=True)
    finally:
        ctx.close()


if __name__ == "__main__":
    main()

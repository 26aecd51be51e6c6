"""Benchmark of the MNAD hot path on one MI355X.

Default workload (the largest single-GPU configuration, C5): batched RTM on
the 8001 x 2001 model (dx = dz = 12.5 m, 30 CPML cells -> 2061 x 8061 padded),
seeded layered + faulted + high-velocity-body synthetic, migration model = a
100 m Gaussian smoothing, Ricker 30 Hz, dt = 1 ms, nt = 12000, the 64-shot
layout of evenly spaced surface sources with a +-2 km split spread; the
source wavefield is kept as the BF16 boundary ring the configuration names.
One step = one shot of the batched job: a forward on the true model
(observed gather) and a forward on the migration model (both in the shared
forward launches), the scattered data, and the adjoint with the imaging
condition (beside the reverse reconstruction of the source wavefield).
Useful updates per step = 3 * N * nt (N padded points); the reconstruction
is overhead and not counted.  --workload c2 runs the C2 single-shot RTM
instead; c3 / c5all / c4 the many-shot and serving forms.

Prints one JSON line (metric, value in Gpt-updates/s, roofline of the step
kernel from HIP events on the library's stream, CPU baseline of the FP64
oracle on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpt-updates/s & RTM shots/hr vs plain FP32 at ≤1e-4 rel err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


# HBM bytes per step-kernel launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes (each scaled by what the counter reads for a 256 MiB device copy:
# x2.0 fetch, x1.0 write), keyed by (workload, step-kernel symbol).  The
# counters cannot be read from inside a timed run, so the line reports the
# figure of the committed counter runs when the same kernel runs the same
# workload, and null otherwise.
MEASURED_TRAFFIC = {
    # C2 RTM shot with f16 snapshots, nt = 300: forward (2 shots) 69.68 MB,
    # adjoint 35.58 MB, equal launch counts -> 52.63 MB per launch
    # (profiles/k_stepw_c2_rtm_shot.md, where the kernel's symbol is
    # k_stepw<1, false>; the same code)
    ("c2", "mnad::k_stepw<1>", "f16"): 52.63e6,
    # C5, 20 shots, BF16 ring, groups of 10, pipelined: 2.0 x 6,480,359 +
    # 5,777,001 KiB per launch (profiles/k_stepx_c5_ring_bf16_batched.md)
    ("c5", "mnad::k_stepx<4>", "ring_bf16", 20, 10, True): 19.187e9,
    # the same with step kernel 13 (16-row segments, 4 waves): 30,835,394 KiB per
    # launch with its passes (profiles/k_step2q_c5_ring_bf16_batched.md)
    ("c5", "mnad::k_step2q<4, true>", "ring_bf16", 20, 10, True): 31.575e9,
}


SNAP_NOTES = {
    "ring_bf16": ("BF16 boundary ring (the configuration's storage): the source wavefield is "
                  "rebuilt by a reverse reconstruction beside the adjoint"),
    "ring_f16": ("FP16 boundary ring: the source wavefield is rebuilt by a reverse "
                 "reconstruction beside the adjoint"),
    "f16": "f16 d2u snapshots with a per-shot power-of-two scale",
    "bf16": ("bf16 d2u snapshots: the same bytes as f16, but the normalized C2 image misses "
             "the 1e-3 bound (2.2e-3)"),
    "f32": "f32 d2u snapshots",
}


def kernel_name(t, dx_eq_dz=True):
    """Device symbol (as rocprofv3 prints it) of the step kernel a tuning runs
    (kernel 13: the instance for grids with dx == dz when dx_eq_dz)."""
    k, r, b = t["kernel"], t["rows_per_thread"], t["block_rows"]
    return {0: f"mnad::k_step<{r}, {b}>", 4: f"mnad::k_stepw<{b}>",
            6: f"mnad::k_stepx<{b}>", 7: f"mnad::k_step2w<{4 * r}, {b}>",
            9: "mnad::k_res", 10: f"mnad::k_stepx2<{b}>", 11: "mnad::k_stepw2",
            13: f"mnad::k_step2q<{b}, {'true' if dx_eq_dz else 'false'}>"}.get(k, f"kernel {k}")


def c2_workload(nt, npml=30):
    from __graft_entry__ import load_package
    load_package()
    import importlib
    synth = importlib.import_module("mnad_amd.synth")
    dx, dt, fc = 10.0, 8e-4, 20.0
    V, Vs = synth.model_c2(seed=2024)
    # column-major, as a MATLAB caller hands them over (no conversion per call)
    v, vs = np.asfortranarray(synth.pad_edges(V, npml)), np.asfortranarray(synth.pad_edges(Vs, npml))
    w = synth.ricker(fc, dt, nt)
    nxp = V.shape[1]
    zs, xs = npml, npml + nxp // 2
    xr = npml + np.arange(nxp)
    zr = np.full(nxp, npml + 1)
    return dict(v=v, vs=vs, w=w, npml=npml, dx=dx, dt=dt, zs=zs, xs=xs, zr=zr, xr=xr)


def c3_workload(nt, nshots):
    """C3: 3401 x 751 (+30 CPML), dx = 10 m, dipping layers, three normal
    faults, a 4500 m/s body and a 200 m water layer (seeded); migration
    model = 100 m Gaussian smoothing; `nshots` sources evenly spaced on the
    first physical row, split-spread receivers +-200 cells (2 km) every cell
    on the second row, clipped at the model edges; Ricker 25 Hz; dt = 0.8 ms."""
    from __graft_entry__ import load_package
    load_package()
    import importlib
    synth = importlib.import_module("mnad_amd.synth")
    npml, dx, dt, fc = 30, 10.0, 8e-4, 25.0
    V, Vs = synth.model_c3(seed=3)
    v, vs = np.asfortranarray(synth.pad_edges(V, npml)), np.asfortranarray(synth.pad_edges(Vs, npml))
    w = synth.ricker(fc, dt, nt)
    nxp = V.shape[1]
    xs_phys = np.rint(np.linspace(0, nxp - 1, nshots)).astype(int)
    recs = []
    for x in xs_phys:
        xr = np.arange(max(0, x - 200), min(nxp, x + 201))
        recs.append((np.full(xr.size, npml + 1), npml + xr))
    return dict(v=v, vs=vs, w=w, npml=npml, dx=dx, dt=dt, zs=[npml] * nshots,
                xs=list(npml + xs_phys), recs=recs)


def c5_workload(nt, nshots, crop=False):
    """C5: 8001 x 2001 (+30 CPML), dx = dz = 12.5 m, faulted layered model with
    a 4500 m/s body (seeded), Ricker 30 Hz, dt = 1 ms.  `nshots` sources evenly
    spaced on the first physical row, split-spread receivers +-160 cells (2 km)
    on the second row.  crop=True: the 1001 x 501 window at the model centre
    with one shot in its middle and all 1001 receivers (the FP64 oracle case)."""
    from __graft_entry__ import load_package
    load_package()
    import importlib
    synth = importlib.import_module("mnad_amd.synth")
    npml, dx, dt, fc = 30, 12.5, 1e-3, 30.0
    V, Vs = synth.model_c5(seed=5)
    if crop:
        z0, x0 = 0, (V.shape[1] - 1001) // 2
        V, Vs = V[z0:z0 + 501, x0:x0 + 1001], Vs[z0:z0 + 501, x0:x0 + 1001]
    v, vs = np.asfortranarray(synth.pad_edges(V, npml)), np.asfortranarray(synth.pad_edges(Vs, npml))
    w = synth.ricker(fc, dt, nt)
    nxp = V.shape[1]
    if crop:
        xs_phys = np.array([nxp // 2])
        recs = [(np.full(nxp, npml + 1), npml + np.arange(nxp))]
    else:
        xs_phys = np.rint(np.linspace(0, nxp - 1, nshots)).astype(int)
        recs = []
        for x in xs_phys:
            xr = np.arange(max(0, x - 160), min(nxp, x + 161))
            recs.append((np.full(xr.size, npml + 1), npml + xr))
    wl = dict(v=v, vs=vs, w=w, npml=npml, dx=dx, dt=dt, zs=[npml] * len(xs_phys),
              xs=list(npml + xs_phys), recs=recs)
    if crop:   # single-shot keys of the C2 path
        wl.update(zs=npml, xs=int(npml + xs_phys[0]), zr=recs[0][0], xr=recs[0][1])
    return wl


def c2_multi_workload(nt, nshots):
    """C2 model with `nshots` sources evenly spaced on the first physical row,
    each with the full 1001-receiver line."""
    wl = c2_workload(nt)
    npml, nxp = wl["npml"], wl["v"].shape[1] - 2 * wl["npml"]
    xs_phys = np.rint(np.linspace(50, nxp - 51, nshots)).astype(int)
    wl.update(zs=[npml] * nshots, xs=list(npml + xs_phys),
              recs=[(wl["zr"], wl["xr"])] * nshots)
    return wl


def run_multi(args, mnad, wl, label):
    """Many-shot RTM: the batched form (autotuned kernel variant and shots per
    launch, hipGraph chunks) against the plain form (one shot and one
    propagation per launch sequence, plain launches) on the same workload."""
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    fmt = mnad.SNAP_FORMATS[args.snap]
    N = v.shape[0] * v.shape[1]
    nshots = len(wl["xs"])
    plain_flags = mnad.FLAG_SERIAL | mnad.FLAG_NO_GRAPH
    with mnad.Context(0) as ctx:
        t0 = time.perf_counter()
        if args.tuning:   # a fixed kernel variant and --batch shots per group (no autotune)
            k, rz, br = (int(x) for x in args.tuning.split(","))
            ctx.set_tuning(kernel=k, rows_per_thread=rz, block_rows=br)
            tuning, per_launch, us = ctx.tuning(), args.batch, 0.0
        else:
            tuning, per_launch, us = ctx.autotune(g, v, max_shots=32)
        tune_s = time.perf_counter() - t0
        print(f"# autotune {tune_s:.2f} s: {tuning} shots/launch {per_launch} ({us:.2f} us/shot-step)",
              file=sys.stderr, flush=True)
        sid = ctx.stack_create(g)
        t0 = time.perf_counter()
        for k in range(args.plain_shots):
            ctx.rtm_shots(sid, g, v, wl["vs"], wl["w"], [wl["zs"][k]], [wl["xs"][k]],
                          [wl["recs"][k]], snap_fmt=fmt, batch=1, flags=plain_flags)
        plain_s = (time.perf_counter() - t0) / max(args.plain_shots, 1)
        print(f"# plain path: {plain_s:.3f} s/shot", file=sys.stderr, flush=True)
        ctx.stack_destroy(sid)
        sid = ctx.stack_create(g)
        ctx.timing(reset=True)
        t0 = time.perf_counter()
        # slices of 16 shots (one or more launch groups each), with a progress line per slice
        for k0 in range(0, nshots, 16):
            k1 = min(nshots, k0 + 16)
            ctx.rtm_shots(sid, g, v, wl["vs"], wl["w"], wl["zs"][k0:k1], wl["xs"][k0:k1],
                          wl["recs"][k0:k1], snap_fmt=fmt, batch=args.batch)
            print(f"# shots {k1}/{nshots} done ({time.perf_counter() - t0:.1f} s)",
                  file=sys.stderr, flush=True)
        el = time.perf_counter() - t0
        k_ms, k_launches, k_bytes = ctx.timing(reset=True)
        img, ill, ns = ctx.stack_image(sid, g)
        ctx.stack_destroy(sid)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    nt = wl["w"].size
    out = {
        "metric": METRIC, "value": round(3.0 * N * nt * nshots / el / 1e9, 3),
        "unit": "Gpt-updates/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(el * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded model)",
        "config": {"workload": label, "grid_padded": [int(v.shape[0]), int(v.shape[1])],
                   "shots": nshots, "nt": nt, "snapshots": args.snap,
                   "tuning": tuning, "shots_per_launch": per_launch},
        "shots_per_hour": round(3600.0 * nshots / el, 1),
        "plain_shots_per_hour": round(3600.0 / plain_s, 1),
        "speedup_vs_plain": round(plain_s * nshots / el, 3),
        "autotune_s": round(tune_s, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": kernel_name(tuning),
                     "avg_launch_us": round(k_ms / max(k_launches, 1) * 1e3, 3),
                     "launches": int(k_launches)},
        "image_finite": bool(np.isfinite(img).all()),
    }
    print(json.dumps(out), flush=True)


def run_c5_headline(args, mnad):
    """The headline line: the C5 batched RTM, one step = one shot of the
    batched job (the reference's parfor over shots, main_Acoustic_RTM.m:
    223-226,250-254, as the batch axis).  The shots are the C5 layout's 64
    sources in order (cycled past 64); --warmup shots run first in one batched
    call, then --steps shots in one timed batched call (autotuned kernel and
    shots per launch, hipGraph chunks).  Storage of the source wavefield:
    --snap (default: the BF16 boundary ring the configuration names).  Extra
    keys: the same job with f16 snapshots, the plain path (one shot and one
    propagation per launch sequence, no graph), the C2 single-shot RTM."""
    nt = 12000 if args.nt == 4000 else args.nt
    wl = c5_workload(nt, 64)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    N = v.shape[0] * v.shape[1]
    fmt = mnad.SNAP_FORMATS[args.snap]
    nsh = len(wl["xs"])

    def pick(k0, n):
        idx = [(k0 + q) % nsh for q in range(n)]
        return ([wl["zs"][q] for q in idx], [wl["xs"][q] for q in idx], [wl["recs"][q] for q in idx])

    def job(ctx, k0, n, sf, batch=0, flags=0):
        sid = ctx.stack_create(g)
        zs, xs, recs = pick(k0, n)
        ctx.rtm_shots(sid, g, v, wl["vs"], wl["w"], zs, xs, recs, snap_fmt=sf, batch=batch,
                      flags=flags)
        img, ill, _ = ctx.stack_image(sid, g)
        ctx.stack_destroy(sid)
        return img, ill

    print(f"# C5 {v.shape[0]}x{v.shape[1]} nt={nt} snap={args.snap}", file=sys.stderr, flush=True)
    extra = {}
    with mnad.Context(0) as ctx:
        t0 = time.perf_counter()
        if args.tuning:   # a fixed kernel variant and --batch shots per group (no autotune)
            k, rz, br = (int(x) for x in args.tuning.split(","))
            ctx.set_tuning(kernel=k, rows_per_thread=rz, block_rows=br)
            tuning, per_launch, us = ctx.tuning(), args.batch, 0.0
        else:
            tuning, per_launch, us = ctx.autotune(g, v, max_shots=32)
        tune_s = time.perf_counter() - t0
        print(f"# tuning {tune_s:.1f} s: {tuning}, {per_launch} shots/launch, "
              f"{us:.2f} us/shot-step", file=sys.stderr, flush=True)
        if args.warmup > 0:
            job(ctx, 0, args.warmup, fmt, batch=args.batch)
            print(f"# warmup: {args.warmup} shots done", file=sys.stderr, flush=True)
        ctx.synchronize()
        ctx.timing(reset=True)
        t0 = time.perf_counter()
        img, ill = job(ctx, args.warmup, args.steps, fmt, batch=args.batch)
        ctx.synchronize()
        el = time.perf_counter() - t0
        k_ms, k_launches, k_bytes = ctx.timing(reset=True)
        group, pipelined = ctx.rtm_schedule()
        print(f"# timed: {args.steps} shots in {el:.1f} s", file=sys.stderr, flush=True)
        if not args.no_plain:
            t1 = time.perf_counter()
            job(ctx, 0, 1, fmt, batch=1, flags=mnad.FLAG_SERIAL | mnad.FLAG_NO_GRAPH)
            plain_s = time.perf_counter() - t1
            extra["plain_ms_per_step"] = round(plain_s * 1e3, 1)
            extra["speedup_vs_plain"] = round(plain_s / (el / args.steps), 3)
            print(f"# plain path: {plain_s:.2f} s per shot", file=sys.stderr, flush=True)
        if args.compare_snap and args.snap != "f16":
            n2 = 4
            t1 = time.perf_counter()
            job(ctx, args.warmup, n2, mnad.SNAP_F16)
            s2 = (time.perf_counter() - t1) / n2
            extra["snapshots_f16"] = {
                "ms_per_step": round(s2 * 1e3, 1),
                "value": round(3.0 * N * nt / s2 / 1e9, 3), "shots": n2,
                "bytes_per_shot_source_wavefield": int(
                    snap_bytes(mnad, g, wl["vs"], nt, mnad.SNAP_F16))}
            print(f"# f16 snapshots: {s2:.2f} s per shot", file=sys.stderr, flush=True)
        extra["bytes_per_shot_source_wavefield"] = int(snap_bytes(mnad, g, wl["vs"], nt, fmt))
        if not args.no_c2:
            extra["c2_single_shot"] = c2_single_shot(args, mnad)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    kname = kernel_name(tuning)
    out = {
        "metric": METRIC, "value": round(3.0 * N * nt * args.steps / el / 1e9, 3),
        "unit": "Gpt-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered + faulted + high-velocity-body model)",
        "config": {"workload": (f"C5: 8001x2001 (+30 CPML) dx=dz=12.5 m, nt={nt}, dt=1 ms, Ricker "
                                "30 Hz, 64-shot layout (split spread +-2 km), batched RTM, one "
                                f"step = one shot; source wavefield: {args.snap}"),
                   "grid_padded": [int(v.shape[0]), int(v.shape[1])], "nt": nt,
                   "source_wavefield": args.snap, "tuning": tuning,
                   "shots_per_launch_tuned": per_launch,
                   "shots_per_group": group,
                   "schedule": "pipelined" if pipelined else "in order",
                   "updates_per_step": 3 * N * nt},
        "shots_per_hour": round(3600.0 * args.steps / el, 1),
        "autotune_s": round(tune_s, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get(("c5", kname, args.snap, args.steps, group,
                                                      pipelined)),
                     "traffic_unit": "bytes per launch (rocprofv3 counter runs)",
                     "algorithmic_bytes_per_launch": round(k_bytes / max(k_launches, 1)),
                     "kernel": kname, "avg_launch_us": round(k_ms / max(k_launches, 1) * 1e3, 3),
                     "launches": int(k_launches)},
        # img = Im ./ illum is 0/0 where no shot of the stack reached (far parts of
        # the 100 km line); it must be finite everywhere else
        "image_finite_where_illuminated": bool(np.isfinite(img[ill > 0]).all()),
        "illuminated_fraction": round(float(np.mean(ill > 0)), 4),
    }
    out.update(extra)
    if not args.no_cpu and args.cpu_steps > 0:
        print("# cpu baseline (FP64 oracle, C5 window) ...", file=sys.stderr, flush=True)
        cw = c5_workload(nt, 1, crop=True)
        out["cpu_baseline"] = cpu_baseline(cw, min(args.cpu_steps, nt), args.cpu_procs,
                                           label="C5 1001x501 window")
    print(json.dumps(out), flush=True)


def snap_bytes(mnad, g, vs, nt, fmt):
    """Device bytes one shot keeps for its adjoint: K d2u snapshots of the
    interior (fwd_PDE_exact_MNAD.m:137-142), or the boundary ring (three
    fields on 4 (nx + nz - 4 npml + 4) cells every step, 16-bit) plus the last
    two levels of the three fields."""
    nzi, nxi = g.nz - 2 * g.npml, g.nx - 2 * g.npml
    _, T, K, _ = mnad.snap_schedule(nt, float(np.min(vs)), float(np.max(vs)))
    if fmt in (mnad.SNAP_RING_F16, mnad.SNAP_RING_BF16):
        ring = 4 * ((nxi + 4) + nzi)
        return nt * 3 * ring * 2 + 6 * g.nz * g.nx * 4
    return K * nzi * nxi * (4 if fmt == mnad.SNAP_F32 else 2)


def c2_single_shot(args, mnad, shots=3):
    """C2 single-shot RTM (the observed-data forward and the migration-model
    forward in one batched launch sequence, then the adjoint): seconds per
    shot with --snap storage and on the plain path."""
    wl = c2_workload(4000)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    N = v.shape[0] * v.shape[1]
    res = {}
    with mnad.Context(0) as ctx:
        for name, fmt in ((args.snap, mnad.SNAP_FORMATS[args.snap]), ("f16", mnad.SNAP_F16)):
            sid = ctx.stack_create(g)
            call = (sid, g, v, wl["vs"], wl["w"], [wl["zs"]], [wl["xs"]], [(wl["zr"], wl["xr"])])
            ctx.rtm_shots(*call, snap_fmt=fmt, batch=1)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(shots):
                ctx.rtm_shots(*call, snap_fmt=fmt, batch=1)
            ctx.synchronize()
            s = (time.perf_counter() - t0) / shots
            t0 = time.perf_counter()
            ctx.rtm_shots(*call, snap_fmt=fmt, batch=1, flags=mnad.FLAG_SERIAL | mnad.FLAG_NO_GRAPH)
            plain = time.perf_counter() - t0
            ctx.stack_destroy(sid)
            res[name] = {"ms_per_shot": round(s * 1e3, 2), "value": round(3.0 * N * 4000 / s / 1e9, 3),
                         "plain_ms_per_shot": round(plain * 1e3, 2),
                         "speedup_vs_plain": round(plain / s, 3)}
            if name == "f16":
                break
    print(f"# C2 single shot: {res}", file=sys.stderr, flush=True)
    return res


class C4Requests:
    """The C4 serving workload's requests, seeded: physical grids
    U[256,2048] x U[256,1024] (dx = 10 m, 20 CPML cells), layered + random
    models 1500-4500 m/s cut from one correlated-noise field, Ricker
    U[10,30] Hz, nt U[1000,8000], dt = 0.8*0.65*dx/vmax
    (main_Acoustic_Data.m:71,77), 70 % modelling / 30 % RTM, Poisson
    arrivals at `qps`.  build(k) returns request k's arrays."""

    npml, dx = 20, 10.0

    def __init__(self, n, seed, qps=10.0):
        from __graft_entry__ import load_package
        load_package()
        import importlib
        self.synth = importlib.import_module("mnad_amd.synth")
        rng = np.random.default_rng(seed)
        npml, dx = self.npml, self.dx
        NZ, NX = 1024 + 2 * npml, 2048 + 2 * npml
        self.noise = self.synth.correlated_noise((NZ, NX), 50.0, dx, rng)
        self.noise_s = self.synth.gaussian_smooth(self.noise, 100.0, dx)
        self.arrivals = np.cumsum(rng.exponential(1.0 / qps, n))
        self.specs = []
        for _ in range(n):
            nxp, nzp = int(rng.integers(256, 2049)), int(rng.integers(256, 1025))
            self.specs.append(dict(nxp=nxp, nzp=nzp, fc=float(rng.uniform(10, 30)),
                                   nt=int(rng.integers(1000, 8001)), rtm=bool(rng.random() < 0.3),
                                   oz=int(rng.integers(0, NZ - nzp - 2 * npml + 1)),
                                   ox=int(rng.integers(0, NX - nxp - 2 * npml + 1)),
                                   xs=int(rng.integers(0, nxp))))

    def build(self, k):
        """(nz, nx, npml, dx, dt, v, vs, w, zs, xs, zr, xr) of request k;
        vs is None for a modelling request."""
        sp, npml, dx = self.specs[k], self.npml, self.dx
        nz, nx = sp["nzp"] + 2 * npml, sp["nxp"] + 2 * npml
        z = np.arange(nz) * dx
        base = 1500.0 + 2700.0 * (z / z[-1])[:, None]
        crop = np.s_[sp["oz"]:sp["oz"] + nz, sp["ox"]:sp["ox"] + nx]
        v = np.asfortranarray(np.clip(base * (1 + 0.03 * self.noise[crop]), 1500.0, 4500.0))
        vs = (np.asfortranarray(np.clip(base * (1 + 0.03 * self.noise_s[crop]), 1500.0, 4500.0))
              if sp["rtm"] else None)
        dt = 0.8 * 0.65 * dx / 4500.0
        w = self.synth.ricker(sp["fc"], dt, sp["nt"])
        xr = npml + np.arange(0, sp["nxp"], 4)
        zr = np.full(xr.size, npml + 1)
        return nz, nx, npml, dx, dt, v, vs, w, npml, npml + sp["xs"], zr, xr


def run_c4(args, mnad):
    """Serving: Poisson arrivals at --qps of the C4 requests (C4Requests).
    Reports p50/p99 latency (submit -> result) and aggregate Gpt-updates/s."""
    import threading
    reqs = C4Requests(args.requests, args.seed, args.qps)
    n = args.requests
    arrivals, specs = reqs.arrivals, reqs.specs

    def build(k):
        nz, nx, npml, dx, dt, v, vs, w, zs, xs, zr, xr = reqs.build(k)
        return mnad.Grid(nz, nx, npml, dx, dx, dt), v, vs, w, zs, xs, zr, xr, nz * nx

    lat, results, errors = {}, {}, []
    tickets = []
    lock = threading.Lock()
    done_evt = threading.Event()
    with mnad.Context(0) as ctx:
        if args.tuning:   # the server's one-step kernel (default: the library's)
            k, rz, br = (int(x) for x in args.tuning.split(","))
            ctx.set_tuning(kernel=k, rows_per_thread=rz, block_rows=br)
        ctx.serve_start(hbm_budget=int(args.budget_gb * 2**30), max_active=args.max_active,
                        chunk_steps=args.chunk, max_points=int(args.max_points))

        def waiter():
            k = 0
            while k < n:
                with lock:
                    ready = k < len(tickets)
                    item = tickets[k] if ready else None
                if not ready:
                    time.sleep(0.0005)
                    continue
                idx, t, ups = item
                try:
                    out, info = ctx.wait(t)
                    lat[idx] = (info["done_s"] - info["submit_s"], info["admit_s"] - info["submit_s"])
                    results[idx] = (info["submit_s"], info["done_s"], ups)
                except Exception as e:   # noqa: BLE001 - reported in the JSON
                    errors.append(f"{idx}: {e}")
                k += 1
                if k % 100 == 0:
                    print(f"# {k}/{n} requests done", file=sys.stderr, flush=True)
            done_evt.set()

        th = threading.Thread(target=waiter, daemon=True)
        th.start()
        t0 = time.perf_counter()
        for k, sp in enumerate(specs):
            g, v, vs, w, zs, xs, zr, xr, N = build(k)
            delay = arrivals[k] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            kind = mnad.REQ_RTM if sp["rtm"] else mnad.REQ_MODEL
            t = ctx.submit(kind, g, v, w, zs, xs, zr, xr, v_smooth=vs, snap_fmt=mnad.SNAP_F16,
                           want_gather=not sp["rtm"])
            ups = N * sp["nt"] * (3 if sp["rtm"] else 1)
            with lock:
                tickets.append((k, t, ups))
        th.join()
        ctx.serve_stop()
    L = np.array([lat[k][0] for k in sorted(lat)]) * 1e3
    Q = np.array([lat[k][1] for k in sorted(lat)]) * 1e3
    first = min(r[0] for r in results.values())
    last = max(r[1] for r in results.values())
    updates = sum(r[2] for r in results.values())
    out = {
        "metric": METRIC, "value": round(updates / (last - first) / 1e9, 3),
        "unit": "Gpt-updates/s", "n_gpus": 1, "steps": n, "warmup": 0,
        "ms_per_step": round((last - first) * 1e3 / n, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered + random models)",
        "config": {"workload": f"C4 serving: {n} requests, Poisson {args.qps}/s, 70% modelling / 30% RTM, "
                               "grids 256x256..2048x1024 (+20 CPML), nt 1000..8000",
                   "qps": args.qps, "requests": n, "chunk_steps": args.chunk,
                   "step_kernel": args.tuning or "4,1,1 (library default)",
                   "budget_gb": args.budget_gb, "max_active": args.max_active,
                   "max_points": args.max_points},
        "latency_ms": {"p50": round(float(np.percentile(L, 50)), 2),
                       "p90": round(float(np.percentile(L, 90)), 2),
                       "p99": round(float(np.percentile(L, 99)), 2),
                       "max": round(float(L.max()), 2)},
        "queue_wait_ms": {"p50": round(float(np.percentile(Q, 50)), 2),
                          "p99": round(float(np.percentile(Q, 99)), 2)},
        "makespan_s": round(last - first, 2), "errors": errors[:5], "n_errors": len(errors),
    }
    print(json.dumps(out), flush=True)


def run_probe(args, mnad):
    """Forward shots (two, or the fourth number of --probe) x 32 steps on the
    workload's grid, run twice in one launch sequence with the given step
    kernel: a short, fixed sequence for hardware-counter runs of one kernel
    variant."""
    p = [int(x) for x in args.probe.split(",")]
    k, rz, br = p[:3]
    ns = p[3] if len(p) > 3 else 2
    if args.workload == "c5":
        wl = c5_workload(32, ns)
    elif args.workload == "c3":
        wl = c3_workload(32, ns)
    else:
        wl = c2_multi_workload(32, ns)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    with mnad.Context(0) as ctx:
        ctx.set_tuning(kernel=k, rows_per_thread=rz, block_rows=br)
        for _ in range(2):
            ctx.fwd_shots(g, v, wl["w"], wl["zs"], wl["xs"], wl["recs"], batch=ns)
        ms, n, b = ctx.timing()
    print(json.dumps({"probe": args.probe, "workload": args.workload,
                      "grid_padded": [int(v.shape[0]), int(v.shape[1])],
                      "us_per_launch": round(ms * 1e3 / max(n, 1), 2), "launches": int(n),
                      "kernel": kernel_name(dict(kernel=k, rows_per_thread=rz, block_rows=br))}),
          flush=True)


def run_compare(args, mnad):
    """Step-kernel time per step (HIP events on the library stream, from
    mnad_timing) of several kernel variants on forward batches: for each
    padded grid of --grid (NZxNXxNPML, comma list), each variant of --compare
    (KERNEL:ROWS:BLOCK_ROWS, comma list) and each shot count of --compare-shots,
    one untimed call (set-up, graph capture) and --compare-calls timed calls of
    fwd_shots with --nt steps, on a linear-gradient model (1500 to 3000 m/s),
    sources at the centre of the first physical row (3 cells apart), a receiver
    on every interior column of the next row.  --compare-calls 1
    --compare-warm 0 makes one call per variant (for profiler runs)."""
    dt, fc = 8e-4, 20.0
    t = np.arange(args.nt) * dt
    tau = np.pi * (t - 1.1 / fc) * fc
    w = (1 - 2 * tau**2) * np.exp(-tau**2)
    with mnad.Context(0) as ctx:
        for gs in args.grid.split(","):
            nz, nx, npml = (int(x) for x in gs.split("x"))
            v = np.asfortranarray(1500.0 + 1500.0 * np.arange(nz)[:, None] / nz * np.ones((1, nx)))
            g = mnad.Grid(nz, nx, npml, 10.0, 10.0, dt)
            xr = np.arange(npml, nx - npml)
            zr = np.full(xr.size, npml + 1)
            for ks in args.compare.split(","):
                k, r, b = (int(x) for x in ks.split(":"))
                ctx.set_tuning(kernel=k, rows_per_thread=r, block_rows=b)
                for ns in (int(x) for x in args.compare_shots.split(",")):
                    xs = [nx // 2 + 3 * q for q in range(ns)]
                    call = (g, v, w, [npml] * ns, xs, [(zr, xr)] * ns)
                    try:
                        for _ in range(args.compare_warm):
                            ctx.fwd_shots(*call, batch=ns)
                        ctx.timing(reset=True)
                        for _ in range(args.compare_calls):
                            ctx.fwd_shots(*call, batch=ns)
                        ms, launches, _ = ctx.timing(reset=True)
                        us = ms * 1e3 / (args.nt * args.compare_calls)
                        print(f"{nz}x{nx} kernel {k},{r},{b} {ns} shot(s): {us:.2f} us per step, "
                              f"{us / ns:.2f} us per shot-step ({launches} launches)", flush=True)
                    except mnad.MnadError as e:
                        print(f"{nz}x{nx} kernel {k},{r},{b} {ns} shot(s): {e}", flush=True)


def _cpu_shot(args):
    """One oracle shot sample in a worker process: (points * steps, seconds)."""
    wl, steps = args
    from oracle import oracle as O
    v = wl["v"]
    g = O.grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    t0 = time.perf_counter()
    O.forward(g, v, wl["w"][:steps], wl["zs"], wl["xs"], wl["zr"], wl["xr"], want_state=False)
    return v.size * steps, time.perf_counter() - t0


def host_cpus():
    """CPUs of the host (nproc) and those this process may run on."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    return os.cpu_count() or 1, aff


def cpu_baseline(wl, steps, procs=0, label=""):
    """FP64 oracle on the first `steps` forward steps of the true-model shot of
    the same workload: once on one core, then one such shot per core on `procs`
    cores at once (the reference's parfor over shots); the aggregate rate is
    the value.  procs = 0: every CPU this process may run on, at most 16 (the
    CPU share a one-GPU job gets on the GPU hosts; nproc is reported beside)."""
    from multiprocessing import get_context
    sample = dict(v=wl["v"], w=wl["w"], npml=wl["npml"], dx=wl["dx"], dt=wl["dt"],
                  zs=wl["zs"], xs=wl["xs"], zr=wl["zr"], xr=wl["xr"])
    n1, t1 = _cpu_shot((sample, steps))
    single = n1 / t1 / 1e9
    nproc, aff = host_cpus()
    procs = procs or min(16, aff)
    if procs > 1:
        with get_context("spawn").Pool(procs) as pool:
            t0 = time.perf_counter()
            res = pool.map(_cpu_shot, [(sample, steps)] * procs)
            el = time.perf_counter() - t0
        value = sum(r[0] for r in res) / el / 1e9
    else:
        value, el = single, t1
    where = f" of the {label}" if label else ""
    return dict(value=value, unit="Gpt-updates/s", cores=procs, kind="port",
                single_core_value=single, host_nproc=nproc, host_cpus_allowed=aff,
                sample=f"FP64 oracle forward{where}, true model, first {steps} of {wl['w'].size} "
                       f"steps; 1 shot on 1 core ({t1:.1f} s), then {procs} such shots on {procs} "
                       f"cores at once ({el:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c5", choices=["c2", "c3", "c4", "c5", "c5all", "c5crop"],
                    help="c5 (default): the C5 batched RTM, one step = one shot; c5all: all 64 "
                         "C5 shots as one step; c2: the C2 single-shot RTM")
    ap.add_argument("--requests", type=int, default=1000, help="C4: number of requests")
    ap.add_argument("--qps", type=float, default=10.0, help="C4: Poisson arrival rate")
    ap.add_argument("--chunk", type=int, default=32, help="C4: steps between admissions")
    ap.add_argument("--budget-gb", type=float, default=0.0, help="C4: HBM budget (0 = 80%% free)")
    ap.add_argument("--max-active", type=int, default=64, help="C4: active requests at most")
    ap.add_argument("--max-points", type=float, default=4e6,
                    help="C4: admit while fewer grid points than this are stepping (0: no limit)")
    ap.add_argument("--seed", type=int, default=4, help="C4: workload seed")
    ap.add_argument("--save-images-only", action="store_true",
                    help="with --save-outputs: image and illumination only (no gathers)")
    ap.add_argument("--save-outputs", default="",
                    help="C2: after timing, run one more shot and save its gathers, image and "
                         "illumination to this .npz (for tests/validate_full_size.py)")
    ap.add_argument("--shots", type=int, default=128, help="C3: number of shots")
    ap.add_argument("--batch", type=int, default=0, help="C3: shots per launch sequence (0 = by memory)")
    ap.add_argument("--plain-shots", type=int, default=2, help="many-shot runs: shots timed on the plain path")
    ap.add_argument("--shots-c2", type=int, default=1,
                    help="C2 grid with this many shots, batched vs plain (default: the 1-shot bench)")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (C5: shots of the batched job; default 20 for C5, "
                         "whose groups of 10 then pipeline, else 5)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed warm-up steps (default 5 for C5, else 1)")
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--snap", default="",
                    choices=["", "f32", "f16", "bf16", "ring_f16", "ring_bf16"],
                    help="source-wavefield storage (default: the configuration's boundary "
                         "storage, ring_bf16 for C2 and C5, ring_f16 for C3)")
    ap.add_argument("--no-compare-snap", dest="compare_snap", action="store_false",
                    help="C5: skip the f16-snapshot comparison")
    ap.add_argument("--no-c2", action="store_true", help="C5: skip the C2 single-shot extra keys")
    ap.add_argument("--npml", type=int, default=30, help="C2: CPML cells (30 in the workload)")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="CPU baseline sample steps (0: 300 on the C5 window, 400 on C2)")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="CPU baseline processes (0 = host cores, at most 16)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--probe", default="",
                    help="KERNEL,ROWS,BLOCK_ROWS[,SHOTS]: time SHOTS (default 2) forward shots "
                         "x 32 steps (twice) of the workload's grid with that step kernel "
                         "(for counter runs)")
    ap.add_argument("--compare", default="",
                    help="KERNEL:ROWS:BLOCK_ROWS list: step-kernel time per step of each variant "
                         "on forward batches of the --grid grids (see run_compare)")
    ap.add_argument("--grid", default="561x1061x30", help="--compare: padded NZxNXxNPML list")
    ap.add_argument("--compare-shots", default="1,2", help="--compare: shots per launch list")
    ap.add_argument("--compare-calls", type=int, default=1, help="--compare: timed calls")
    ap.add_argument("--compare-warm", type=int, default=1, help="--compare: untimed calls")
    ap.add_argument("--tuning", default="",
                    help="step kernel K,RZ,BR for the timed shots, without the autotuner "
                         "(batched forms: --batch shots per group; default: autotuned, or the "
                         "library's kernel for the single shot)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (x3) to calibrate the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="only time plain 1 and 4 GiB device copies and fills (HIP events): "
                         "the HBM rate a simple streaming kernel reaches on this device")
    args = ap.parse_args()
    if args.copy_ceiling:
        import torch
        out = {}
        for gib in (1, 4):
            n = gib * 2**30 // 4
            x = torch.ones(n, device="cuda")
            y = torch.empty_like(x)
            for name, fn, nbytes in (("copy", lambda: y.copy_(x), 8 * n),
                                     ("fill", lambda: y.fill_(2.0), 4 * n)):
                for _ in range(3):
                    fn()
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(10):
                    fn()
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / 10
                out[f"{name}_{gib}GiB_GBps"] = round(nbytes / ms / 1e6, 1)
            del x, y
        print(json.dumps(out), flush=True)
        return
    if args.copy_calibration:
        import torch
        n = 256 * 2**20 // 4
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by "
              "shots with one image all-reduce, which this single-GPU build does not run.",
              file=sys.stderr)
        sys.exit(2)

    if args.steps is None:
        args.steps = 20 if args.workload == "c5" else 5
    if args.warmup is None:
        args.warmup = 5 if args.workload == "c5" else 1
    from __graft_entry__ import load_package
    mnad = load_package()
    if args.probe:
        run_probe(args, mnad)
        return
    if args.compare:
        if args.nt == 4000:
            args.nt = 1000
        run_compare(args, mnad)
        return
    if args.workload == "c4":
        run_c4(args, mnad)
        return
    if not args.snap:   # the boundary storage each configuration names
        args.snap = "ring_f16" if args.workload == "c3" else "ring_bf16"
    if args.cpu_steps == 0:
        args.cpu_steps = 300 if args.workload == "c5" else 400
    if args.workload == "c5":
        run_c5_headline(args, mnad)
        return
    if args.workload == "c5all":
        nt = 12000 if args.nt == 4000 else args.nt
        run_multi(args, mnad, c5_workload(nt, args.shots),
                  f"C5: 8001x2001 (+30 CPML) dx=12.5 m, {args.shots} shots batched RTM, nt={nt}")
        return
    if args.workload == "c3":
        nt = 6000 if args.nt == 4000 else args.nt
        run_multi(args, mnad, c3_workload(nt, args.shots),
                  f"C3: 3401x751 (+30 CPML), {args.shots} shots batched RTM, nt={nt}")
        return
    if args.shots_c2 > 1:
        run_multi(args, mnad, c2_multi_workload(args.nt, args.shots_c2),
                  f"C2 grid: 1001x501 (+30 CPML), {args.shots_c2} shots batched RTM, nt={args.nt}")
        return
    if args.workload == "c5crop":
        args.nt = 12000 if args.nt == 4000 else args.nt
        wl = c5_workload(args.nt, 1, crop=True)
    else:
        wl = c2_workload(args.nt, args.npml)
    v = wl["v"]
    g = mnad.Grid(v.shape[0], v.shape[1], wl["npml"], wl["dx"], wl["dx"], wl["dt"])
    fmt = mnad.SNAP_FORMATS[args.snap]
    recs = [(wl["zr"], wl["xr"])]
    N = v.shape[0] * v.shape[1]
    nt = args.nt
    print(f"# C2 {v.shape[0]}x{v.shape[1]} nt={nt} snap={args.snap}", file=sys.stderr, flush=True)

    with mnad.Context(0) as ctx:
        if args.tuning:
            k, rz, br = (int(x) for x in args.tuning.split(","))
            ctx.set_tuning(kernel=k, rows_per_thread=rz, block_rows=br)
        sid = ctx.stack_create(g)
        kname = kernel_name(ctx.tuning())

        def shot(flags=0):
            ctx.rtm_shots(sid, g, v, wl["vs"], wl["w"], [wl["zs"]], [wl["xs"]], recs,
                          snap_fmt=fmt, batch=1, flags=flags)

        for k in range(args.warmup):
            shot()
            print(f"# warmup {k + 1}/{args.warmup} done", file=sys.stderr, flush=True)
        ctx.timing(reset=True)
        t0 = time.perf_counter()
        for k in range(args.steps):
            shot()
            print(f"# step {k + 1}/{args.steps} done", file=sys.stderr, flush=True)
        el = time.perf_counter() - t0
        k_ms, k_launches, k_bytes = ctx.timing(reset=True)
        plain_ms = None
        if not args.no_plain:
            t1 = time.perf_counter()
            shot(flags=mnad.FLAG_SERIAL | mnad.FLAG_NO_GRAPH)
            plain_ms = (time.perf_counter() - t1) * 1e3
            print("# plain shot (one propagation per launch sequence, no graph) done",
                  file=sys.stderr, flush=True)
        ctx.stack_destroy(sid)
        if args.save_outputs:
            sid = ctx.stack_create(g)
            dobs, dsyn = ctx.rtm_shots(sid, g, v, wl["vs"], wl["w"], [wl["zs"]], [wl["xs"]], recs,
                                       snap_fmt=fmt, batch=1, want_gathers=True)
            img, ill, _ = ctx.stack_image(sid, g)
            ctx.stack_destroy(sid)
            extra = {} if args.save_images_only else dict(dobs=dobs[0].astype(np.float32),
                                                          dsyn=dsyn[0].astype(np.float32))
            np.savez_compressed(args.save_outputs, img=img, illum=ill, nt=nt, snap=args.snap,
                                **extra)
            print(f"# saved {args.save_outputs}", file=sys.stderr, flush=True)

    ms_per_step = el / args.steps * 1e3
    updates = 3.0 * N * nt * args.steps
    value = updates / el / 1e9
    avg_launch_us = k_ms / max(k_launches, 1) * 1e3
    achieved = k_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered + correlated-random model)",
        "config": {"workload": ("C5 window: 1001x501 (+30 CPML) of the C5 model, dx=12.5 m, "
                                f"nt={nt}, dt=1 ms, Ricker 30 Hz, 1001 receivers; modelling + "
                                "1-shot RTM per step") if args.workload == "c5crop" else
                               ("C2: 1001x501 (+30 CPML) dx=10 m, nt=4000, dt=0.8 ms, Ricker "
                                "20 Hz, 1001 receivers; modelling + 1-shot RTM per step"),
                   "grid_padded": [int(v.shape[0]), int(v.shape[1])], "nt": nt,
                   "snapshots": args.snap, "updates_per_step": 3 * N * nt,
                   "snapshots_note": SNAP_NOTES.get(args.snap, "")},
        "shots_per_hour": round(3600.0 / (ms_per_step / 1e3), 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((args.workload, kname, args.snap)),
                     "traffic_unit": "bytes per launch (rocprofv3 counter runs)",
                     "algorithmic_bytes_per_launch": round(k_bytes / max(k_launches, 1)),
                     "kernel": kname, "avg_launch_us": round(avg_launch_us, 3),
                     "launches": int(k_launches)},
    }
    if plain_ms is not None:
        out["plain_ms_per_step"] = round(plain_ms, 3)
        out["speedup_vs_plain"] = round(plain_ms / ms_per_step, 3)
    if not args.no_cpu and args.cpu_steps > 0:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(wl, min(args.cpu_steps, nt), args.cpu_procs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark: forward+adjoint gradient throughput on one MI355X.

A step is one full gradient evaluation of the batched P-SV Marmousi-shaped
workload (737x240 cells at 12.5 m, 16 shots, 368 receivers, dt 1 ms,
nt 4000): forward of every shot with snapshots, waveform misfit and adjoint
sources, adjoint with lambda/mu/rho kernel accumulation, one-sided Gaussian
smoothing of the three kernels.  All 16 shots run as one batch per launch.
Inputs (model, observed data) are resident in HBM before the timed region.

Prints one JSON line (rank 0).  Extra keys: roofline (stencil step vs HBM
peak, timed live with HIP events inside the library), cpu_baseline (the CPU
oracle on a bounded sample of the same workload), plain (the one-shot-per-
launch GPU path on the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gcell-updates/s fwd+adjoint; gradients/hr; p50/p99 shot latency vs plain"
HBM_PEAK_GBPS = 8000.0
# Algorithmic HBM bytes per cell-update of one P-SV stencil time step: each of
# the 7 wavefields (vx, vz, ux, uz, sxx, szz, sxz) read and written once and
# the 3 material fields read once, FP32.
PSV_STEP_BYTES = 4 * (2 * 7 + 3)
# HBM bytes per cell of one launch of the default step kernel (fused column
# kernel, tile 58x8, fast math, P-SV), measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE and scaled by a 256 MiB copy (fetch x2.0, write x1.0); see
# profiles/k_col_psv_marmousi737x240_16shots.md.  Only reported for that
# configuration.
MEASURED_TRAFFIC_PER_CELL = 69.33
PROFILED_TILE = (0, 8, 1, 1, 16)  # variant, rows, columns per lane, math, batch


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16, help="shots per launch")
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--sigma", type=int, default=5, help="Gaussian filter sigma (cells)")
    ap.add_argument("--prof-every", type=int, default=16,
                    help="time every n-th step's stencil launches with HIP events (0 = off)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-shot GPU path")
    ap.add_argument("--cpu-shots", type=int, default=1)
    ap.add_argument("--variant", type=int, default=0, help="0 fused step kernel, 1 kernel pair")
    ap.add_argument("--tile-x", type=int, default=0,
                    help="fused tile rows (0: inv_autotune picks rows and columns per lane)")
    ap.add_argument("--cols", type=int, default=1, help="columns per lane with --tile-x")
    ap.add_argument("--math", type=int, default=1, help="0 exact (IEEE divisions), 1 fast")
    ap.add_argument("--graphs", type=int, default=1, help="capture time loops in HIP graphs")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=12,
                    help="mixed-size serving leg: requests per run (0 = skip)")
    ap.add_argument("--serve-rate-x", type=float, default=1.0,
                    help="serving leg arrival rate, in multiples of the plain loop's capacity")
    ap.add_argument("--serving-sweep", type=int, default=0,
                    help="run only the serving sweep over this many requests (one JSON line "
                         "per run) and exit")
    return ap.parse_args()


def serving_leg(inv, n, rate_x, snap_format):
    """p50/p99 latency and throughput of the request server against the plain
    one-request-at-a-time loop on one Poisson arrival sequence of n mixed-size
    requests (grids 256x256 .. 2048x1024, nt 1000-6000, SH/P-SV, forward or
    forward+adjoint, 1-8 shots), at rate_x times the plain loop's capacity."""
    S = inv.serving
    log = lambda m: print(m, file=sys.stderr, flush=True)
    reqs = S.make_requests(n)
    kw = dict(snap_format=snap_format, max_lanes=4, max_slots=16)
    base, _ = S.run_load(reqs, rate=None, plain=True, log=log, **kw)
    rate = rate_x * base["requests_per_s"]
    plain, _ = S.run_load(reqs, rate=rate, plain=True, log=log, **kw)
    batched, _ = S.run_load(reqs, rate=rate, plain=False, log=log, **kw)
    keys = ("p50_latency_s", "p99_latency_s", "requests_per_s", "gcell_per_s", "batches",
            "mean_requests_per_batch", "peak_bytes")
    return {"requests": n, "rate_per_s": round(rate, 4), "rate_x_plain_capacity": rate_x,
            "plain_capacity_req_per_s": round(base["requests_per_s"], 4),
            "snapshots": ["fp32", "fp16", "bf16"][snap_format],
            "plain": {k: plain[k] for k in keys}, "batched": {k: batched[k] for k in keys},
            "p99_speedup": round(plain["p99_latency_s"] / batched["p99_latency_s"], 3)}


def configure_kernel(ctx, a, count):
    """The step kernel: autotuned (default) or as given; returns its tile."""
    if a.variant == 0 and a.tile_x == 0:
        v, tx, cols = ctx.autotune(0, count, steps=20)
    elif a.variant == 0:
        ctx.set_tile(a.tile_x, a.cols)
        v, tx, cols = 0, a.tile_x, a.cols
    else:
        ctx.set_variant(a.variant, a.tile_x)
        v, tx, cols = a.variant, a.tile_x, 1
    return {"variant": v, "rows": tx, "cols_per_lane": cols}


def cpu_baseline(P, shots):
    """The CPU oracle (FP32, OpenMP) on `shots` shots of the same workload."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle  # the checker, used here only as the timed CPU baseline

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    pyoracle.set_threads(threads)
    Ps = P.with_(src_ix=P.src_ix[:shots], src_iz=P.src_iz[:shots], stf_x=P.stf_x[:shots],
                 stf_y=P.stf_y[:shots], stf_z=P.stf_z[:shots])
    ab = pyoracle.absbound(P.nx, P.nz, P.abs_width, P.abs_alpha, P.abs_left, P.abs_right,
                           P.abs_bottom, P.abs_top)
    O = pyoracle.Problem(P.nx, P.nz, P.dx, P.dz, P.dt, P.nt, P.lam0, P.mu0, P.rho0, ab,
                         Ps.src_ix, Ps.src_iz, Ps.stf_x, Ps.stf_y, Ps.stf_z, P.rec_ix, P.rec_iz,
                         sh=P.sh, psv=P.psv, sfe=P.sfe, adj=1)
    obs = np.zeros((shots, P.ncomp, P.nrec, P.nt), np.float32)
    t0 = time.perf_counter()
    O.gradient(obs, 0, P.sigma, dtype=np.float32)
    dt = time.perf_counter() - t0
    work = Ps.cell_updates(shots)
    return {"value": work / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP32 gradient (forward+misfit+adjoint+filter) of {shots} of the "
                      f"16 shots, full nt={P.nt}, {P.nx}x{P.nz}; {dt:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU "
              "shot sharding is described in DESIGN.md but not built.", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as graft

    inv = graft.load_package()
    if a.serving_sweep:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        for s in inv.serving.sweep(a.serving_sweep, log=log,
                                   snap_format=inv.invhip.INV_SNAP_F16):
            print(json.dumps(s), flush=True)
        return
    P = inv.problems.marmousi(nt=a.nt, sigma=a.sigma)
    B = min(a.batch, P.nsrc)
    print(f"# workload {P.name} {P.nx}x{P.nz} shots={P.nsrc} nt={P.nt} batch={B}",
          file=sys.stderr, flush=True)
    ctx = P.context(batch=B, model="true")
    ctx.set_math(a.math)
    ctx.set_graphs(a.graphs)
    tile = configure_kernel(ctx, a, B)
    # observed data: forward on the true model (Misfit::init, misfit.h:135-148)
    for first in range(0, P.nsrc, B):
        n = min(B, P.nsrc - first)
        ctx.forward(first, n, snapshots=False)
        ctx.obs_from_traces(first, n)
    P.load_model(ctx, "init")
    ctx.sync()
    print("# observed data ready", file=sys.stderr, flush=True)

    if a.copy_calib_mb:
        nb = a.copy_calib_mb << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        ctx.vec_fill(src, 1.0, nb // 4)
        ctx.vec_copy(dst, src, nb // 4)
        ctx.sync()
        ctx.free(src)
        ctx.free(dst)
    for _ in range(a.warmup):
        ctx.gradient(P.misfit, P.sigma)
    ctx.sync()
    ctx.prof_reset()
    ctx.prof_enable(a.prof_every)
    t0 = time.perf_counter()
    f = 0.0
    for _ in range(a.steps):
        f = ctx.gradient(P.misfit, P.sigma)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.prof_enable(0)
    ms = (t1 - t0) / a.steps * 1e3
    work = P.cell_updates()
    value = work / (ms / 1e3) / 1e9
    print(f"# timed {a.steps} gradients: {ms:.1f} ms each, misfit {f:.6e}", file=sys.stderr,
          flush=True)

    nv, tv = ctx.prof_read(inv.invhip.INV_PROF_VELOCITY)
    ns, ts = ctx.prof_read(inv.invhip.INV_PROF_STRESS)
    nf, tf = ctx.prof_read(inv.invhip.INV_PROF_STEP)
    nfr, tfr = ctx.prof_read(inv.invhip.INV_PROF_FWD_RUN)
    nar, tar = ctx.prof_read(inv.invhip.INV_PROF_ADJ_RUN)
    roof = None
    cells = B * P.N
    step_s = None
    if nfr:
        # replayed forward graphs: every launch in them but a handful of
        # set-up fills is the step kernel, so time / steps is its duration
        step_s = tfr / nfr / 1e3
        kernel = "fused P-SV time step (k_col / k_colv, autotuned tile), from forward-graph replays"
        avg = {"step": round(step_s * 1e6, 2)}
        samples = {"forward_steps": nfr}
        if nar:
            avg["adjoint_step_incl_interaction"] = round(tar / nar * 1e3, 2)
    elif nf:
        step_s = tf / nf / 1e3
        kernel = "fused P-SV time step (k_col / k_colv)"
        avg = {"step": round(tf / nf * 1e3, 2)}
        samples = {"step": nf}
    elif nv and ns:
        step_s = (tv / nv + ts / ns) / 1e3
        kernel = "P-SV stencil step (velocity + stress launch pair)"
        avg = {"velocity": round(tv / nv * 1e3, 2), "stress": round(ts / ns * 1e3, 2)}
        samples = {"velocity": nv, "stress": ns}
    if step_s:
        achieved = PSV_STEP_BYTES * cells / step_s / 1e9
        profiled = ((tile["variant"], tile["rows"], tile["cols_per_lane"], a.math, B) ==
                    PROFILED_TILE and nfr)
        traffic = round(MEASURED_TRAFFIC_PER_CELL * cells) if profiled else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": kernel, "alg_bytes_per_launch": PSV_STEP_BYTES * cells,
                "avg_us": avg, "samples": samples}
    ni, ti = ctx.prof_read(inv.invhip.INV_PROF_INTERACT)
    ctx.close()

    plain = None
    if not a.no_plain:
        pctx = P.context(batch=1, model="true")
        configure_kernel(pctx, a, 1)
        pctx.set_math(a.math)
        pctx.set_graphs(False)  # the plain path: one shot per launch, no graph
        for s in range(P.nsrc):
            pctx.forward(s, 1, snapshots=False)
            pctx.obs_from_traces(s, 1)
        P.load_model(pctx, "init")
        pctx.gradient(P.misfit, P.sigma)
        pctx.sync()
        t0 = time.perf_counter()
        pctx.gradient(P.misfit, P.sigma)
        pctx.sync()
        pms = (time.perf_counter() - t0) * 1e3
        pctx.close()
        plain = {"ms_per_gradient": round(pms, 2), "gcell_per_s": round(work / pms / 1e6, 2),
                 "gradients_per_hr": round(3.6e6 / pms, 1),
                 "note": "one shot per launch, same kernels, no batching, no graph"}
        print(f"# plain path: {pms:.1f} ms per gradient", file=sys.stderr, flush=True)

    cpu = None if a.no_cpu else cpu_baseline(P, a.cpu_shots)
    serving = None
    if a.serve_requests > 0:
        serving = serving_leg(inv, a.serve_requests, a.serve_rate_x, inv.invhip.INV_SNAP_F16)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Marmousi-shaped model; observed data modelled on the true model)",
        "config": {"workload": "batched P-SV gradient, Marmousi-shaped 737x240 @ 12.5 m, "
                               "16 shots, 368 receivers, dt 1 ms, nt 4000",
                   "nx": P.nx, "nz": P.nz, "shots": P.nsrc, "receivers": P.nrec, "nt": P.nt,
                   "batch": B, "sfe": P.sfe, "filter_sigma": P.sigma, "misfit": "waveform",
                   "snapshots": "fp32 in HBM", "step_kernel_tile": tile},
        "gradients_per_hr": round(3.6e6 / ms, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain": plain,
        "serving": serving,
    }
    if ni:
        out["interaction_avg_us"] = round(ti / ni * 1e3, 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

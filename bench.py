#!/usr/bin/env python3
"""Benchmark: forward+adjoint gradient throughput on one MI355X.

A step is one full gradient evaluation (Misfit::run(kernel=true),
misfit/misfit.h:15-51): forward of every shot with snapshots, misfit and
adjoint sources, adjoint with kernel accumulation, one-sided Gaussian
smoothing of the kernels.  Inputs (model, observed data) are resident in HBM
before the timed region.  Workloads (--workload):

  c0            P-SV checkerboard 200x200 at 10 m (+-10 % squares of 20 cells
                on vp, vs, rho), 4 explosive Ricker sources at 20 cells depth,
                40 receivers, dt 1 ms, nt 2000, waveform misfit, sigma 5; all 4
                shots per launch; the CPU oracle runs the whole gradient on the
                same observed data and the kernels are compared with the GPU's.
  c1 (default)  batched P-SV, Marmousi-shaped 737x240 at 12.5 m, 16 shots,
                368 receivers, dt 1 ms, nt 4000, waveform misfit, FP32
                snapshots; all 16 shots per launch.
  c2            SH, the same generator at 1474x480 and 6.25 m, 64 shots,
                737 receivers, dt 0.5 ms, nt 8000, FP16 snapshots every 10
                steps with 4000-step checkpoints (the adjoint recomputes the first
                half of the forward; one half's FP16 snapshots take 72 GB);
                plus the error of FP32 / FP16 / BF16 snapshot storage against
                the FP64 oracle on a 2-shot, 2000-step sample.
  c4            P-SV, the generator at 4096x2048 and 5 m, dt 0.4 ms,
                nt 10000, 2048 receivers, envelope misfit, FP16 snapshots with
                1000-step checkpoints; timed on 16 of the 256 shots (one
                launch; a full gradient is 16 such batches), plus the FP64
                check on the 1024x512 / 20 m downsampled model.

Prints one JSON line (rank 0).  Extra keys: roofline (stencil step vs HBM
peak, timed live with HIP events inside the library), cpu_baseline (the CPU
oracle on a bounded sample of the same workload; c0: all of it), parity (c0:
GPU gradient against the CPU oracle's), plain (c1: the one-shot-
per-launch GPU path on the same workload), serving (c1: p50/p99 latency of
the request server against the plain loop), accuracy (c2, c4).
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
# SH: the 4 wavefields (vy, uy, sxy, szy) read and written, mu and rho read
SH_STEP_BYTES = 4 * (2 * 4 + 2)
# HBM bytes per cell of one launch, measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE (per dispatch, each scaled by what the counter reads for a plain
# device copy of known size: fetch x2.0, write x1.0), for the tiles the
# autotuner picks on these workloads, fast math, the forward variants
# without snapshot roles (the launches the live timing samples most);
# keyed by (workload, steps per launch, rows, columns per lane, shots per
# launch):
#   c1 one-step 4 x 2, 16 shots:  68.30 B (algorithmic 68),
#                                 profiles/k_colv_psv_marmousi737x240_16shots.md
#   c4 two-step 20 x 2, 4 shots:  71.51 B per launch of two steps (algorithmic 68),
#                                 profiles/k_pair_psv_4096x2048_4shots.md
#   c2 two-step 16 x 2, 64 shots: 40.36 B per launch of two steps (algorithmic 40),
#                                 profiles/k_pair_sh_marmousi1474x480_64shots.md
MEASURED_TRAFFIC_PER_CELL = {("c1", 1, 4, 2, 16): 68.30, ("c4", 2, 20, 2, 4): 71.51,
                             ("c2", 2, 16, 2, 64): 40.36}


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0, help="shots per launch (0: the workload's)")
    ap.add_argument("--ckpt", type=int, default=-1,
                    help="checkpoint interval in steps for 16-bit snapshots (-1: the workload's, "
                         "0: none, every snapshot of the run resident)")
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c4"])
    ap.add_argument("--nt", type=int, default=0, help="time steps (0: the workload's)")
    ap.add_argument("--sigma", type=int, default=5, help="Gaussian filter sigma (cells)")
    ap.add_argument("--prof-every", type=int, default=16,
                    help="time every n-th step's stencil launches with HIP events (0 = off)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-shot GPU path")
    ap.add_argument("--no-accuracy", action="store_true", help="c2/c4: skip the FP64 check")
    ap.add_argument("--cpu-shots", type=int, default=8,
                    help="c1: shots of the CPU oracle's timed sample")
    ap.add_argument("--variant", type=int, default=0, help="0 fused step kernel, 1 kernel pair")
    ap.add_argument("--tile-x", type=int, default=0,
                    help="fused tile rows (0: inv_autotune picks rows and columns per lane)")
    ap.add_argument("--cols", type=int, default=1, help="columns per lane with --tile-x")
    ap.add_argument("--pair-rows", type=int, default=-1,
                    help="two-step kernel tile rows (-1: the autotuner decides, 0: one step "
                         "per launch)")
    ap.add_argument("--pair-cols", type=int, default=2, help="two-step kernel columns per lane")
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
    return ap.parse_args(argv)


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


def configure_kernel(ctx, a, count, pair=True):
    """The step kernels: autotuned (default) or as given; returns the tiles.
    pair=False keeps one step per launch."""
    if a.variant == 0 and a.tile_x == 0:
        # a one-step tile and, over it, a two-step tile if one is faster
        v, tx, cols = ctx.autotune(0, count, steps=20)
    elif a.variant == 0:
        ctx.set_tile(a.tile_x, a.cols)
        v, tx, cols = 0, a.tile_x, a.cols
    else:
        ctx.set_variant(a.variant, a.tile_x)
        v, tx, cols = a.variant, a.tile_x, 1
    if not pair or a.variant != 0 or (a.tile_x != 0 and a.pair_rows < 0):
        ctx.set_pair(0)
    elif a.pair_rows >= 0:
        ctx.set_pair(a.pair_rows, a.pair_cols)
    tile = {"variant": v, "rows": tx, "cols_per_lane": cols}
    pr, pv = ctx.get_pair()
    if pr:
        tile["two_step"] = {"rows": pr, "cols_per_lane": pv}
    return tile


def cpu_baseline(P, shots, nt=None, obs=None):
    """The CPU oracle (FP32, OpenMP) on `shots` shots of the same workload
    (the first nt steps when nt is given), against observed data `obs`
    ([shots][ncomp][nrec][nt], zeros when None: the work is the same).
    Returns the baseline record and the oracle's (misfit, k_lambda, k_mu,
    k_rho)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle  # the checker, used here only as the timed CPU baseline

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    pyoracle.set_threads(threads)
    nt = nt or P.nt
    Ps = P.with_(src_ix=P.src_ix[:shots], src_iz=P.src_iz[:shots],
                 stf_x=np.ascontiguousarray(P.stf_x[:shots, :nt]),
                 stf_y=np.ascontiguousarray(P.stf_y[:shots, :nt]),
                 stf_z=np.ascontiguousarray(P.stf_z[:shots, :nt]), nt=nt,
                 src_type=None if P.src_type is None else P.src_type[:shots])
    ab = pyoracle.absbound(P.nx, P.nz, P.abs_width, P.abs_alpha, P.abs_left, P.abs_right,
                           P.abs_bottom, P.abs_top)
    O = pyoracle.Problem(P.nx, P.nz, P.dx, P.dz, P.dt, nt, P.lam0, P.mu0, P.rho0, ab,
                         Ps.src_ix, Ps.src_iz, Ps.stf_x, Ps.stf_y, Ps.stf_z, P.rec_ix, P.rec_iz,
                         sh=P.sh, psv=P.psv, sfe=P.sfe, adj=1, src_type=Ps.src_type)
    if obs is None:
        obs = np.zeros((shots, P.ncomp, P.nrec, nt), np.float32)
    t0 = time.perf_counter()
    res = O.gradient(obs, P.misfit, P.sigma, dtype=np.float32)
    dt = time.perf_counter() - t0
    work = Ps.cell_updates(shots)
    return {"value": work / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP32 gradient (forward+misfit+adjoint+filter) of {shots} of the "
                      f"{P.nsrc} shots, nt={nt} of {P.nt}, {P.nx}x{P.nz}; {dt:.1f} s"}, res


def workload(inv, a):
    """(problem, shots per launch, snapshot format, checkpoint interval,
    bytes per cell-update of the step kernel, description) of a workload."""
    pr, H = inv.problems, inv.invhip
    if a.workload == "c0":
        P = pr.checker(nt=a.nt or 2000, sigma=a.sigma, explosive=True)
        return (P, 4, H.INV_SNAP_F32, 0, PSV_STEP_BYTES,
                f"batched P-SV gradient, checkerboard 200x200 @ 10 m, 4 explosive shots, "
                f"40 receivers, dt 1 ms, nt {P.nt}, sigma {P.sigma}")
    if a.workload == "c1":
        P = pr.marmousi(nt=a.nt or 4000, sigma=a.sigma)
        return (P, 16, H.INV_SNAP_F32, 0, PSV_STEP_BYTES,
                f"batched P-SV gradient, Marmousi-shaped 737x240 @ 12.5 m, 16 shots, "
                f"368 receivers, dt 1 ms, nt {P.nt}")
    if a.workload == "c2":
        # SH waves do not enter the water layer (mu = 0 there): shots and
        # receivers sit on the seafloor, 2 rows below the 12 water rows
        P = pr.marmousi(nx=1474, nz=480, dx=6.25, dt=5e-4, nt=a.nt or 8000, nsrc=64, nrec=737,
                        f0=8.0, t0=0.15, seed=2, sh=1, psv=0, sigma=a.sigma, src_depth=14,
                        rec_depth=14, name="marmousi-sh")
        # every FP16 snapshot of the 64 shots stays in HBM (145 GB): no recomputation
        return (P, 64, H.INV_SNAP_F16, 0, SH_STEP_BYTES,
                f"batched SH gradient, Marmousi-shaped 1474x480 @ 6.25 m, 64 shots, "
                f"737 seafloor receivers, {{batch}} per launch, dt 0.5 ms, nt {P.nt}, FP16 "
                f"snapshots every 10 steps, {{ckpt}}")
    P = pr.marmousi(nx=4096, nz=2048, dx=5.0, dt=4e-4, nt=a.nt or 10000, nsrc=256, nrec=2048,
                    f0=12.0, t0=0.1, seed=4, sigma=a.sigma, name="marmousi-large")
    P = P.with_(misfit=1)
    pick = np.arange(0, P.nsrc, 16)  # 16 of the 256 shots, spread over the line
    P = P.with_(src_ix=P.src_ix[pick], src_iz=P.src_iz[pick], stf_x=P.stf_x[pick],
                stf_y=P.stf_y[pick], stf_z=P.stf_z[pick])
    # four shots per launch keep all of their FP16 snapshots in HBM (268 GB):
    # no recomputation.  Sixteen per launch would need 1000-step checkpoints
    # and recompute nine tenths of the forward (39.8 s against 29.0 s)
    return (P, 4, H.INV_SNAP_F16, 0, PSV_STEP_BYTES,
            f"batched P-SV gradient, 4096x2048 @ 5 m, 16 of 256 shots, {{batch}} per launch, "
            f"2048 receivers, dt 0.4 ms, nt {P.nt}, envelope misfit, FP16 snapshots every 10 "
            f"steps, {{ckpt}}")


def accuracy(inv, a):
    """Gradient and misfit error of FP32, FP16-snapshot and BF16-snapshot
    storage (all arithmetic FP32) against the FP64 oracle, on a sample the
    CPU runs in seconds: c2 on its own model (2 shots, nt 2000), c4 on the
    1024x512 / 20 m downsampled model (2 shots, dt 2 ms, nt 2000)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle  # the checker: the FP64 reference of this comparison

    pr, H = inv.problems, inv.invhip
    if a.workload == "c2":
        P = pr.marmousi(nx=1474, nz=480, dx=6.25, dt=5e-4, nt=2000, nsrc=2, nrec=737, f0=8.0,
                        t0=0.15, seed=2, sh=1, psv=0, sigma=0, src_depth=14, rec_depth=14,
                        name="marmousi-sh")
        ckpt, sample = 400, "1474x480 SH, 2 shots, nt 2000 of 8000"
    else:
        P = pr.marmousi(nx=1024, nz=512, dx=20.0, dt=2e-3, nt=2000, nsrc=2, nrec=512, f0=6.0,
                        t0=0.2, seed=4, sigma=0, name="marmousi-large-ds").with_(misfit=1)
        ckpt, sample = 400, "1024x512 @ 20 m P-SV (c4 downsampled), 2 shots, dt 2 ms, nt 2000"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    pyoracle.set_threads(threads)
    ctx = P.context(batch=2, model="true")
    ctx.forward(0, 2, snapshots=False)
    obs = np.stack([ctx.traces_get(b) for b in range(2)])
    ctx.close()
    ab = pyoracle.absbound(P.nx, P.nz, P.abs_width, P.abs_alpha, P.abs_left, P.abs_right,
                           P.abs_bottom, P.abs_top)
    O = pyoracle.Problem(P.nx, P.nz, P.dx, P.dz, P.dt, P.nt, P.lam0, P.mu0, P.rho0, ab,
                         P.src_ix, P.src_iz, P.stf_x, P.stf_y, P.stf_z, P.rec_ix, P.rec_iz,
                         sh=P.sh, psv=P.psv, sfe=P.sfe, adj=1)
    t0 = time.perf_counter()
    f64, *k64 = O.gradient(obs.astype(np.float64), P.misfit, 0, dtype=np.float64)
    tcpu = time.perf_counter() - t0
    names = ("k_lambda", "k_mu", "k_rho")
    out = {"sample": sample, "reference": f"oracle FP64 ({tcpu:.1f} s on {threads} threads)",
           "misfit_fp64": f64}
    for label, fmt, ck in (("fp32", H.INV_SNAP_F32, 0), ("fp16", H.INV_SNAP_F16, ckpt),
                           ("bf16", H.INV_SNAP_BF16, ckpt)):
        ctx = P.context(batch=2, model="init")
        ctx.set_snapshots(fmt, ck)
        for s_ in range(2):
            ctx.obs_set(s_, obs[s_])
        f = ctx.gradient(misfit_type=P.misfit, sigma=0)
        ks = ctx.kernels_get()
        ctx.close()
        row = {"misfit_rel_err": abs(f - f64) / abs(f64), "checkpoints": ck}
        # whole grid, and the solid cells only (mu > 0): in the water layer the
        # mu kernel is singular next to the sources and FP32 itself (GPU and
        # CPU alike) departs from FP64 there
        solid = P.mu0 > 0
        for n, x, y in zip(names, ks, k64):
            if np.any(y):
                row[n + "_rel_l2"] = float(np.linalg.norm(x - y) / np.linalg.norm(y))
                row[n + "_rel_l2_solid"] = float(np.linalg.norm((x - y)[solid]) /
                                                 np.linalg.norm(y[solid]))
        out[label] = row
        print(f"# accuracy {label}: {row}", file=sys.stderr, flush=True)
    return out


def launch_plan(a, P, B, ckpt, desc):
    """Shots per launch, checkpoint interval and the workload description
    after the --batch / --ckpt overrides."""
    if a.batch > 0:
        B = min(a.batch, P.nsrc)
    if a.ckpt >= 0:
        ckpt = a.ckpt
    if ckpt and P.nt % ckpt:
        ckpt = P.nt  # a shortened run (--nt) the interval does not divide: one segment
    desc = desc.format(batch=B, ckpt=f"{ckpt}-step checkpoints" if ckpt else "no checkpoints")
    return B, ckpt, desc


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): sharding the "
              "shots over several GPUs is designed but not built.", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as graft

    inv = graft.load_package()
    if a.serving_sweep:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        for s in inv.serving.sweep(a.serving_sweep, log=log,
                                   snap_format=inv.invhip.INV_SNAP_F16):
            print(json.dumps(s), flush=True)
        return
    P, B, snap_fmt, ckpt, step_bytes, desc = workload(inv, a)
    B, ckpt, desc = launch_plan(a, P, B, ckpt, desc)
    print(f"# workload {a.workload} {P.name} {P.nx}x{P.nz} shots={P.nsrc} nt={P.nt} batch={B}",
          file=sys.stderr, flush=True)
    ctx = P.context(batch=B, model="true")
    ctx.set_math(a.math)
    ctx.set_graphs(a.graphs)
    ctx.set_snapshots(snap_fmt, ckpt)  # before any run sizes the snapshot storage
    tile = configure_kernel(ctx, a, B)
    print(f"# step kernel tile {tile}", file=sys.stderr, flush=True)
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
        print("# warm-up gradient done", file=sys.stderr, flush=True)
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

    H = inv.invhip
    nv, tv = ctx.prof_read(H.INV_PROF_VELOCITY)
    ns, ts = ctx.prof_read(H.INV_PROF_STRESS)
    nf, tf = ctx.prof_read(H.INV_PROF_STEP)
    npr, tpr = ctx.prof_read(H.INV_PROF_PAIR)
    nfr, tfr = ctx.prof_read(H.INV_PROF_FWD_RUN)
    nar, tar = ctx.prof_read(H.INV_PROF_ADJ_RUN)
    ni, ti = ctx.prof_read(H.INV_PROF_INTERACT)
    roof = None
    cells = B * P.N
    two = "two_step" in tile
    waves = "SH" if P.sh and not P.psv else "P-SV" if P.psv and not P.sh else "SH+P-SV"
    # a forward graph holds one kind of step launch when every pair of steps
    # is paired (no 16-bit snapshot scales, no checkpoints, nt even) or none
    uniform = not two or (snap_fmt == H.INV_SNAP_F32 and ckpt == 0 and P.nt % 2 == 0)
    launch_s, kernel, avg, samples, per_launch_steps = None, None, {}, {}, 1
    if nfr and uniform:
        # replayed forward graphs: every launch in them but a handful of
        # set-up fills is the step kernel, so time / launches is its duration
        per_launch_steps = 2 if two else 1
        launch_s = tfr / (nfr / per_launch_steps) / 1e3
        kernel = (f"two-step {waves} kernel k_pair, from forward-graph replays" if two else
                  f"fused {waves} time step (k_col / k_colv, autotuned tile), from forward-graph "
                  "replays")
        avg["two_step_launch" if two else "step"] = round(launch_s * 1e6, 2)
        samples = {"forward_steps": nfr}
    elif two:
        # mixed one- and two-step launches: time each launch of a short direct
        # forward (graphs off) right after the timed region
        if not (npr and nf):
            ctx.set_graphs(False)
            ctx.set_step_limit(min(P.nt, 400))
            ctx.prof_reset()
            ctx.prof_enable(1)
            ctx.forward(0, B, snapshots=True)
            ctx.sync()
            ctx.prof_enable(0)
            ctx.set_step_limit(0)
            nf, tf = ctx.prof_read(H.INV_PROF_STEP)
            npr, tpr = ctx.prof_read(H.INV_PROF_PAIR)
        if npr:
            per_launch_steps = 2
            launch_s = tpr / npr / 1e3
            kernel = f"two-step {waves} kernel k_pair, HIP events on direct launches"
            avg["two_step_launch"] = round(launch_s * 1e6, 2)
            if nf:
                avg["one_step_launch"] = round(tf / nf * 1e3, 2)
            samples = {"two_step": npr, "one_step": nf}
    elif nf:
        launch_s = tf / nf / 1e3
        kernel = f"fused {waves} time step (k_col / k_colv)"
        avg["step"] = round(launch_s * 1e6, 2)
        samples = {"step": nf}
    elif nv and ns:
        launch_s = (tv / nv + ts / ns) / 1e3
        kernel = f"{waves} stencil step (velocity + stress launch pair)"
        avg = {"velocity": round(tv / nv * 1e3, 2), "stress": round(ts / ns * 1e3, 2)}
        samples = {"velocity": nv, "stress": ns}
    if nfr:
        avg["forward_per_step"] = round(tfr / nfr * 1e3, 2)
    if nar:
        avg["adjoint_per_step_incl_interaction"] = round(tar / nar * 1e3, 2)
    if launch_s:
        # each wavefield and model field is read and written once per launch
        achieved = step_bytes * cells / launch_s / 1e9
        t = tile.get("two_step") or tile
        key = (a.workload, per_launch_steps, t["rows"], t["cols_per_lane"], B)
        measured = (MEASURED_TRAFFIC_PER_CELL.get(key)
                    if tile["variant"] == 0 and a.math == 1 else None)
        traffic = round(measured * cells) if measured else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": kernel, "alg_bytes_per_launch": step_bytes * cells,
                "steps_per_launch": per_launch_steps, "avg_us": avg, "samples": samples}
        if per_launch_steps == 2:
            # the same cell-updates at one step per launch would need this
            roof["one_step_equivalent_gbps"] = round(2 * achieved, 1)
    gpu_grad = None
    if a.workload == "c0":
        # the last timed gradient and its observed data, for the CPU comparison
        gpu_grad = (f, ctx.kernels_get(), np.stack([ctx.obs_get(s) for s in range(P.nsrc)]))
    ctx.close()

    plain = None
    if not a.no_plain and a.workload == "c1":
        pctx = P.context(batch=1, model="true")
        configure_kernel(pctx, a, 1, pair=False)
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

    cpu = parity = None
    if not a.no_cpu:
        # c0 in full; c1 in full on a sample of its shots; c2 / c4 on 8 shots x
        # 2000 steps / 2 shots x 100 steps (10-30 s of CPU work each)
        if a.workload == "c0":
            cpu, (fo, *ko) = cpu_baseline(P, P.nsrc, obs=gpu_grad[2])
            fg, kg, _ = gpu_grad
            parity = {"reference": "oracle FP32 gradient of all shots on the same observed data",
                      "math": ["exact", "fast"][a.math],
                      "misfit_rel_err": abs(fg - fo) / abs(fo)}
            for n, x, y in zip(("k_lambda", "k_mu", "k_rho"), kg, ko):
                parity[n + "_rel_l2"] = float(np.linalg.norm(x - y) / np.linalg.norm(y))
            print(f"# parity {parity}", file=sys.stderr, flush=True)
        elif a.workload == "c1":
            cpu = cpu_baseline(P, min(a.cpu_shots, P.nsrc))[0]
        elif a.workload == "c2":
            cpu = cpu_baseline(P, 8, nt=2000)[0]
        else:
            cpu = cpu_baseline(P, 2, nt=100)[0]
    acc = accuracy(inv, a) if a.workload in ("c2", "c4") and not a.no_accuracy else None
    serving = None
    if a.serve_requests > 0 and a.workload == "c1":
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
        "data": ("synthetic (checkerboard model; observed data modelled on the true model)"
                 if a.workload == "c0" else "synthetic (seeded Marmousi-shaped model; observed "
                                            "data modelled on the true model)"),
        "config": {"workload": desc, "name": a.workload,
                   "nx": P.nx, "nz": P.nz, "shots": P.nsrc, "receivers": P.nrec, "nt": P.nt,
                   "batch": B, "sfe": P.sfe, "filter_sigma": P.sigma,
                   "misfit": ["waveform", "envelope"][P.misfit],
                   "snapshots": ["fp32", "fp16", "bf16"][snap_fmt] + " in HBM",
                   "checkpoint_interval": ckpt, "step_kernel_tile": tile},
        "gradients_per_hr": round(3.6e6 / ms, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain": plain,
        "serving": serving,
        "accuracy": acc,
    }
    if parity:
        out["parity"] = parity
    if ni:
        out["interaction_avg_us"] = round(ti / ni * 1e3, 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

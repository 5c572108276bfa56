#!/usr/bin/env python3
"""Benchmark: forward+adjoint gradient throughput on one MI355X.

A step is one full gradient evaluation (Misfit::run(kernel=true),
misfit/misfit.h:15-51) of one wave of shots: forward with snapshots, misfit
and adjoint sources, adjoint with kernel accumulation, one-sided Gaussian
smoothing of the kernels.  Inputs (model, observed data) are resident in HBM
before the timed region.  The workloads are defined in
inv.cu_amd/problems.py (`workload`):

  c4 (default)  P-SV, Marmousi-shaped 4096x2048 at 5 m, 256 shots, 2048
                receivers, dt 0.4 ms, nt 10000, envelope misfit, I16
                snapshots (16-bit fixed point per row segment); a step is the gradient of one wave of 4 shots (one
                launch keeps all of their snapshots in HBM, 268 GB), sigma 5.
                value = Gcell-updates/s of the step; the 256-shot gradient is
                64 such waves (gradients_per_hr_256_shots).
  c0            P-SV checkerboard 200x200, 4 explosive shots, nt 2000: the
                whole gradient is a step; the CPU oracle runs the same
                gradient and the kernels are compared.
  c1            P-SV 737x240, 16 shots in one launch, nt 4000, FP32
                snapshots; also times the one-shot-per-launch plain path and
                10 NLCG iterations through the command-line driver.
  c2            SH 1474x480, 64 shots in one launch, nt 8000, I16
                snapshots all resident (145 GB).

With the default workload, c1 and c2 are also timed (2 gradients each) and
reported under "extra_workloads" (--extras '' skips them).

Prints one JSON line.  Extra keys: roofline (the dominant stencil kernel
against the HBM peak, timed live with HIP events on the library's stream, and
against the bandwidth a plain device copy of 1 GiB reaches on the same GPU),
cpu_baseline (the CPU oracle on a bounded sample of the same workload, on the
threads OMP_NUM_THREADS allows, plus a one-thread sample), accuracy (c2, c4:
gradient error of each snapshot format against the FP64 oracle), serving
(p50/p99 latency of the request server against the plain loop on mixed-size
requests), plain (c1), optimizer_c1 (c1's NLCG run).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gcell-updates/s fwd+adjoint; gradients/hr; p50/p99 shot latency vs plain"
HBM_PEAK_GBPS = 8000.0
# Algorithmic HBM bytes per cell of one stencil launch: each wavefield read and
# written once, each material field read once, FP32.  P-SV: vx, vz, ux, uz,
# sxx, szz, sxz and lambda, mu, rho; SH: vy, uy, sxy, szy and mu, rho.  A
# two-step launch moves them once for two time steps.
PSV_STEP_BYTES = 4 * (2 * 7 + 3)
SH_STEP_BYTES = 4 * (2 * 4 + 2)
# HBM bytes per cell of one launch from rocprofv3 FETCH_SIZE and WRITE_SIZE
# (per dispatch; fetch scaled by 2.0 and write by 1.0, the factors the same
# counters give for a plain device copy of known size), for the tiles the
# autotuner picks on these workloads, fast math, the forward variant without
# snapshot roles; keyed by (workload, steps per launch, rows, columns per lane,
# shots per launch):
#   c1 one-step 4 x 2, 16 shots:  68.30 B (algorithmic 68),
#                                 profiles/k_colv_psv_marmousi737x240_16shots.md
#   c4 two-step 20 x 2, 4 shots:  70.75 B per launch of two steps (algorithmic 68),
#                                 profiles/k_pair_psv_4096x2048_4shots.md
#   c2 two-step 20 x 2, 64 shots: 40.37 B per launch of two steps (algorithmic 40),
#                                 profiles/k_pair_sh_marmousi1474x480_64shots.md
#   c4 three-step 12 x 2, 4 shots: 73.85 B per launch of three steps (algorithmic 68),
#                                 profiles/k_trip_psv_4096x2048_4shots.md
#   c2 three-step 16 x 2, 64 shots: 40.44 B per launch of three steps (algorithmic 40),
#                                 profiles/k_trip_sh_marmousi1474x480_64shots.md
#   c4 walking three-step, chunks of 128 rows, 4 shots: 75.69 B per launch of three
#                                 steps (algorithmic 68), profiles/k_walk_psv_4096x2048_4shots.md
#   c2 walking three-step, chunks of 128 rows, 64 shots: 46.71 B per launch of three
#                                 steps (algorithmic 40), profiles/k_walk_sh_marmousi1474x480_64shots.md
MEASURED_TRAFFIC_PER_CELL = {("c1", 1, 4, 2, 16): 68.30, ("c4", 2, 20, 2, 4): 70.75,
                             ("c2", 2, 20, 2, 64): 40.37,
                             ("c4", 3, 12, 2, 4): 73.85, ("c2", 3, 16, 2, 64): 40.44,
                             ("c4", 3, "walk128", 2, 4): 75.69, ("c2", 3, "walk128", 2, 64): 46.71}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8,
                    help="timed steps (0 with --full-gradient or --lbfgs: only those legs)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4", choices=["c0", "c1", "c2", "c4"])
    ap.add_argument("--extras", default=None,
                    help="comma-separated workloads also timed (2 gradients each); default "
                         "'c1,c2' with --workload c4, none otherwise")
    ap.add_argument("--batch", type=int, default=0, help="shots per launch (0: the workload's)")
    ap.add_argument("--wave", type=int, default=0,
                    help="shots per step (0: c4 one launch, the others all shots)")
    ap.add_argument("--ckpt", type=int, default=-1,
                    help="checkpoint interval in steps for 16-bit snapshots (-1: the workload's, "
                         "0: none, every snapshot of the run resident)")
    ap.add_argument("--snap", default="", choices=["", "fp32", "fp16", "bf16", "i16"],
                    help="snapshot storage (default: the workload's)")
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
    ap.add_argument("--persistent", type=int, default=-1,
                    help="persistent two-step kernel: 1 on (two-step tile 8 x 4), 0 off, -1 as "
                         "the autotuner chose")
    ap.add_argument("--triple", type=int, default=-1,
                    help="three-step tile rows (8, 12, 16 or 20; 0 off; -1 as the autotuner chose)")
    ap.add_argument("--walk", type=int, default=-1,
                    help="walking three-step kernel chunk rows (0 off; -1 as the autotuner "
                         "chose)")
    ap.add_argument("--math", type=int, default=1, help="0 exact (IEEE divisions), 1 fast")
    ap.add_argument("--graphs", type=int, default=1, help="capture time loops in HIP graphs")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="mixed-size serving leg: requests per run (0 = skip)")
    ap.add_argument("--serve-rate-x", type=float, default=4.0,
                    help="serving leg arrival rate, in multiples of the plain loop's capacity")
    ap.add_argument("--full-gradient", action="store_true",
                    help="with the 4096x2048 workload: also time one whole gradient of all 256 "
                         "shots through the command-line driver (about 10 minutes)")
    ap.add_argument("--lbfgs", type=int, default=0,
                    help="with the 4096x2048 workload: also run this many L-BFGS iterations "
                         "over all 256 shots through the command-line driver (about 12 minutes "
                         "for one)")
    ap.add_argument("--lbfgs-shots", type=int, default=256,
                    help="--lbfgs over every (256 / n)-th shot only (n of the 256)")
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
            "snapshots": ["fp32", "fp16", "bf16", "i16"][snap_format],
            "plain": {k: plain[k] for k in keys}, "batched": {k: batched[k] for k in keys},
            "p50_speedup": round(plain["p50_latency_s"] / batched["p50_latency_s"], 3),
            "p99_speedup": round(plain["p99_latency_s"] / batched["p99_latency_s"], 3)}


def configure_kernel(ctx, a, count, nt, pair=True):
    """The step kernels: autotuned (default, on up to 30 steps of the nt of
    the run) or as given; returns the tiles.  pair=False keeps one step per
    launch."""
    if a.variant == 0 and a.tile_x == 0:
        # a one-step tile and, over it, a two-step tile if one is faster
        v, tx, cols = ctx.autotune(0, count, steps=min(30, nt))
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
    if a.persistent == 1:
        ctx.set_pair(8, 4)
    if a.persistent >= 0:
        ctx.set_persistent(a.persistent)
    if a.triple >= 0 and pair and a.variant == 0:
        ctx.set_triple(a.triple)
    if a.walk >= 0 and pair and a.variant == 0:
        ctx.set_walk(a.walk)
    tile = {"variant": v, "rows": tx, "cols_per_lane": cols}
    pr, pv = ctx.get_pair()
    if pr:
        tile["two_step"] = {"rows": pr, "cols_per_lane": pv}
        if ctx.get_persistent():
            tile["two_step"]["persistent"] = True
    if ctx.get_walk():
        tile["three_step"] = {"walk_chunk_rows": ctx.get_walk(), "cols_per_lane": 2}
    elif ctx.get_triple():
        tile["three_step"] = {"rows": ctx.get_triple(), "cols_per_lane": 2}
    return tile


def cpu_info():
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model or platform.processor(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_baseline(P, shots, nt=None, obs=None, sigma=None, threads=None):
    """The CPU oracle (FP32, OpenMP) on `shots` shots of the same workload
    (the first nt steps when nt is given), against observed data `obs`
    ([shots][ncomp][nrec][nt], zeros when None: the work is the same).
    Returns the baseline record and the oracle's (misfit, k_lambda, k_mu,
    k_rho)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle  # the checker, used here only as the timed CPU baseline

    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    pyoracle.set_threads(threads)
    nt = nt or P.nt
    sigma = P.sigma if sigma is None else sigma
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
    res = O.gradient(obs, P.misfit, sigma, dtype=np.float32)
    dt = time.perf_counter() - t0
    work = Ps.cell_updates(shots)
    what = "forward+misfit+adjoint" + (f"+filter (sigma {sigma})" if sigma else "")
    return {"value": work / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP32 {what} of {shots} of the {P.nsrc} shots, nt={nt} of "
                      f"{P.nt}, {P.nx}x{P.nz}, {threads} thread(s); {dt:.1f} s"}, res


def accuracy(inv, name):
    """Gradient and misfit error of FP32, FP16-, BF16- and I16-snapshot
    storage (all arithmetic FP32) against the FP64 oracle: c2 on its own model
    and nt (2 shots), c4 on the 1024x512 / 20 m downsampled model (2 shots, dt
    2 ms, nt 2000).  The kernel error is the relative L2 over the solid cells
    (mu > 0): in the water layer the mu kernel is singular next to the sources
    and FP32 itself (GPU and CPU alike) departs from FP64 there."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle  # the checker: the FP64 reference of this comparison

    pr, H = inv.problems, inv.invhip
    if name == "c2":
        P = pr.shots(pr.workload("c2", sigma=0)[0], [20, 44])
        sample = "1474x480 SH, 2 shots, nt 8000 (full)"
    else:
        P = pr.marmousi(nx=1024, nz=512, dx=20.0, dt=2e-3, nt=2000, nsrc=2, nrec=512, f0=6.0,
                        t0=0.2, seed=4, sigma=0, name="marmousi-large-ds").with_(misfit=1)
        sample = "1024x512 @ 20 m P-SV (c4 downsampled), 2 shots, dt 2 ms, nt 2000"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
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
           "measure": "relative L2 over the solid cells (mu > 0)", "misfit_fp64": f64}
    solid = P.mu0 > 0
    for label, fmt in (("fp32", H.INV_SNAP_F32), ("fp16", H.INV_SNAP_F16),
                       ("bf16", H.INV_SNAP_BF16), ("i16", H.INV_SNAP_I16)):
        ctx = P.context(batch=2, model="init")
        ctx.set_snapshots(fmt, 0)
        for s_ in range(2):
            ctx.obs_set(s_, obs[s_])
        f = ctx.gradient(misfit_type=P.misfit, sigma=0)
        ks = ctx.kernels_get()
        ctx.close()
        row = {"misfit_rel_err": abs(f - f64) / abs(f64)}
        for n, x, y in zip(names, ks, k64):
            if np.any(y):
                row[n] = float(np.linalg.norm((x - y)[solid]) / np.linalg.norm(y[solid]))
        out[label] = row
        log(f"# accuracy {name} {label}: {row}")
    return out


def launch_plan(a, name, P, B, fmt, ckpt, wave):
    """Shots per launch, shots per step, snapshot format and checkpoint
    interval after the command-line overrides."""
    H = sys.modules["inv_cu_amd"].invhip
    if a.batch > 0:
        B = min(a.batch, P.nsrc)
    if a.wave > 0:
        wave = a.wave
    wave = min(max(wave, B), P.nsrc)
    if a.ckpt >= 0:
        ckpt = a.ckpt
    if a.snap:
        fmt = {"fp32": H.INV_SNAP_F32, "fp16": H.INV_SNAP_F16, "bf16": H.INV_SNAP_BF16,
               "i16": H.INV_SNAP_I16}[a.snap]
    if ckpt and P.nt % ckpt:
        ckpt = P.nt  # a shortened run (--nt) the interval does not divide: one segment
    return B, wave, fmt, ckpt


def describe(name, P, B, wave, fmt, ckpt):
    snap = ["FP32", "FP16", "BF16", "I16"][fmt]
    ck = f"{ckpt}-step checkpoints" if ckpt else "every snapshot resident"
    waves = "SH" if P.sh and not P.psv else "P-SV"
    return (f"{name}: batched {waves} gradient, {P.name} {P.nx}x{P.nz} @ {float(P.dx):g} m, "
            f"{wave} of {P.nsrc} shots per step, {B} per launch, {P.nrec} receivers, "
            f"dt {float(P.dt) * 1e3:g} ms, nt {P.nt}, {['waveform', 'envelope'][P.misfit]} "
            f"misfit, sigma {P.sigma}, {snap} snapshots every {P.sfe} steps, {ck}")


def roofline(ctx, inv, a, name, P, B, tile, fmt, ckpt, step_bytes):
    """The dominant stencil kernel's achieved bandwidth: average launch time
    from HIP events on the library's stream, algorithmic bytes per launch."""
    H = inv.invhip
    nv, tv = ctx.prof_read(H.INV_PROF_VELOCITY)
    ns, ts = ctx.prof_read(H.INV_PROF_STRESS)
    nf, tf = ctx.prof_read(H.INV_PROF_STEP)
    npr, tpr = ctx.prof_read(H.INV_PROF_PAIR)
    nfr, tfr = ctx.prof_read(H.INV_PROF_FWD_RUN)
    nar, tar = ctx.prof_read(H.INV_PROF_ADJ_RUN)
    cells = B * P.N
    two = "two_step" in tile
    waves = "SH" if P.sh and not P.psv else "P-SV" if P.psv and not P.sh else "SH+P-SV"
    # a forward graph holds one kind of step launch when every pair of steps
    # is paired (no 16-bit snapshot scales, no checkpoints, nt even) or none
    uniform = not two or (fmt == H.INV_SNAP_F32 and ckpt == 0 and P.nt % 2 == 0
                          and "three_step" not in tile)
    launch_s, kernel, avg, samples, per_launch_steps = None, None, {}, {}, 1
    if nfr and uniform:
        per_launch_steps = 2 if two else 1
        launch_s = tfr / (nfr / per_launch_steps) / 1e3
        res = two and tile["two_step"].get("persistent")
        kernel = (f"persistent two-step {waves} kernel k_res (one launch per forward), from "
                  "forward-graph replays" if res else
                  f"two-step {waves} kernel k_pair, from forward-graph replays" if two else
                  f"fused {waves} time step (k_col / k_colv, autotuned tile), from forward-graph "
                  "replays")
        avg["two_step_launch" if two else "step"] = round(launch_s * 1e6, 2)
        samples = {"forward_steps": nfr}
    elif two:
        # mixed one- and two-step launches: time each launch of a short direct
        # forward (graphs off) right after the timed region
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
        ntr, ttr = ctx.prof_read(H.INV_PROF_TRIPLE)
        if npr:
            avg["two_step_launch"] = round(tpr / npr * 1e3, 2)
        if nf:
            avg["one_step_launch"] = round(tf / nf * 1e3, 2)
        if ntr:
            avg["three_step_launch"] = round(ttr / ntr * 1e3, 2)
        samples = {"two_step": npr, "one_step": nf, "three_step": ntr}
        if ntr and ttr >= tpr:
            # three-step launches take most of the time: they are the
            # dominant kernel (68 B per cell cross HBM once per three steps)
            per_launch_steps = 3
            launch_s = ttr / ntr / 1e3
            kname = "k_walk (walking)" if "walk_chunk_rows" in tile.get("three_step", {}) \
                else "k_trip"
            kernel = f"three-step {waves} kernel {kname}, HIP events on direct launches"
        elif npr:
            per_launch_steps = 2
            launch_s = tpr / npr / 1e3
            kernel = f"two-step {waves} kernel k_pair, HIP events on direct launches"
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
    if not launch_s:
        return None
    achieved = step_bytes * cells / launch_s / 1e9
    t = (tile.get("three_step") if per_launch_steps == 3 else None) or tile.get("two_step") or tile
    key = (name, per_launch_steps, t.get("rows", f"walk{t.get('walk_chunk_rows')}"),
           t["cols_per_lane"], B)
    measured = (MEASURED_TRAFFIC_PER_CELL.get(key)
                if tile["variant"] == 0 and a.math == 1 else None)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": round(measured * cells) if measured else None,
            "kernel": kernel, "alg_bytes_per_launch": step_bytes * cells,
            "steps_per_launch": per_launch_steps, "avg_us": avg, "samples": samples}
    if per_launch_steps > 1:
        # the same cell-updates at one step per launch would need this
        roof["one_step_equivalent_gbps"] = round(per_launch_steps * achieved, 1)
    ni, ti = ctx.prof_read(H.INV_PROF_INTERACT)
    if ni:
        roof["interaction_avg_us"] = round(ti / ni * 1e3, 2)
    return roof


def run_workload(inv, a, name, steps, warmup, headline):
    """Times `steps` gradients of one shot wave of workload `name`; returns
    its record (and, for c0, the GPU gradient for the CPU comparison)."""
    pr, H = inv.problems, inv.invhip
    nt = a.nt if headline else 0
    P, B, fmt, ckpt = pr.workload(name, nt=nt, sigma=a.sigma)
    wave = B if name == "c4" else P.nsrc
    if headline:
        B, wave, fmt, ckpt = launch_plan(a, name, P, B, fmt, ckpt, wave)
    if wave < P.nsrc:
        # a fixed wave of shots spread over the line
        P = pr.shots(P, np.linspace(0, P.nsrc - 1, wave).round().astype(int))
    full_shots = pr.workload(name, nt=nt, sigma=a.sigma)[0].nsrc
    desc = describe(name, P, B, wave, fmt, ckpt)
    log(f"# workload {desc}")
    ctx = P.context(batch=B, model="true")
    ctx.set_math(a.math)
    ctx.set_graphs(a.graphs)
    ctx.set_snapshots(fmt, ckpt)  # before any run sizes the snapshot storage
    tile = configure_kernel(ctx, a, B, P.nt) if headline else configure_kernel(
        ctx, parse(["--math", str(a.math)]), B, P.nt)
    log(f"# step kernel tile {tile}")
    # observed data: forward on the true model (Misfit::init, misfit.h:135-148)
    for first in range(0, P.nsrc, B):
        n = min(B, P.nsrc - first)
        ctx.forward(first, n, snapshots=False)
        ctx.obs_from_traces(first, n)
    P.load_model(ctx, "init")
    ctx.sync()
    log("# observed data ready")
    if headline and a.copy_calib_mb:
        nb = a.copy_calib_mb << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        ctx.vec_fill(src, 1.0, nb // 4)
        ctx.vec_copy(dst, src, nb // 4)
        ctx.sync()
        ctx.free(src)
        ctx.free(dst)
    for _ in range(warmup):
        ctx.gradient(P.misfit, P.sigma)
        log("# warm-up gradient done")
    ctx.sync()
    ctx.prof_reset()
    ctx.prof_enable(a.prof_every)
    t0 = time.perf_counter()
    f = 0.0
    for _ in range(steps):
        f = ctx.gradient(P.misfit, P.sigma)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.prof_enable(0)
    ms = (t1 - t0) / steps * 1e3
    work = P.cell_updates()
    value = work / (ms / 1e3) / 1e9
    log(f"# {name}: timed {steps} gradients of {P.nsrc} shots: {ms:.1f} ms each, "
        f"{value:.1f} Gcell/s, misfit {f:.6e}")
    step_bytes = SH_STEP_BYTES if P.sh and not P.psv else PSV_STEP_BYTES
    roof = roofline(ctx, inv, a, name, P, B, tile, fmt, ckpt, step_bytes)
    gpu_grad = None
    if name == "c0":
        gpu_grad = (f, ctx.kernels_get(), np.stack([ctx.obs_get(s) for s in range(P.nsrc)]))
    ctx.close()
    waves_per_gradient = full_shots / P.nsrc
    rec = {"value": round(value, 3), "ms_per_step": round(ms, 3), "steps": steps,
           "warmup": warmup, "config": {
               "workload": desc, "name": name, "nx": P.nx, "nz": P.nz,
               "shots_per_step": P.nsrc, "shots": full_shots, "receivers": P.nrec, "nt": P.nt,
               "batch": B, "sfe": P.sfe, "filter_sigma": P.sigma,
               "misfit": ["waveform", "envelope"][P.misfit],
               "snapshots": ["fp32", "fp16", "bf16", "i16"][fmt] + " in HBM",
               "checkpoint_interval": ckpt, "step_kernel_tile": tile},
           "roofline": roof}
    if waves_per_gradient > 1:
        rec[f"gradients_per_hr_{full_shots}_shots"] = round(
            3.6e6 / (ms * waves_per_gradient), 2)
        rec["note"] = (f"a step is the gradient of {P.nsrc} of the {full_shots} shots; the "
                       f"{full_shots}-shot gradient is {waves_per_gradient:g} such steps "
                       "(per-shot work is the same), one filter per step")
    else:
        rec["gradients_per_hr"] = round(3.6e6 / ms, 1)
    return rec, P, gpu_grad


def copy_bandwidth(inv, mib=1024, reps=10):
    """The device's practical read+write bandwidth: reps back-to-back device
    copies of `mib` MiB (the library's vec_copy, 16-byte vectors) between two
    synchronisations, best of three such runs; bytes read plus written per
    second."""
    P = inv.problems.tiny(nt=10, nsrc=1, nrec=1)
    ctx = P.context(batch=1, adj=0)
    nb = mib << 20
    src, dst = ctx.alloc(nb), ctx.alloc(nb)
    ctx.vec_fill(src, 1.0, nb // 4)
    ctx.vec_copy(dst, src, nb // 4)
    ctx.sync()
    best = 1e30
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.vec_copy(dst, src, nb // 4)
        ctx.sync()
        best = min(best, (time.perf_counter() - t0) / reps)
    ctx.free(src)
    ctx.free(dst)
    ctx.close()
    return 2 * nb / best / 1e9


def plain_c1(inv, a):
    """The one-shot-per-launch GPU path (no batching, no graph, one step per
    launch) on c1: the plain comparison point of the batched path."""
    P = inv.problems.workload("c1", sigma=a.sigma)[0]
    pctx = P.context(batch=1, model="true")
    configure_kernel(pctx, parse(["--math", str(a.math)]), 1, P.nt, pair=False)
    pctx.set_math(a.math)
    pctx.set_graphs(False)
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
    log(f"# plain path: {pms:.1f} ms per gradient")
    return {"ms_per_gradient": round(pms, 2),
            "gcell_per_s": round(P.cell_updates() / pms / 1e6, 2),
            "gradients_per_hr": round(3.6e6 / pms, 1),
            "note": "c1, one shot per launch, one step per launch, no graph"}


def optimizer_c1(inv, a, iterations=10):
    """c1's inversion: `iterations` NLCG iterations (Optimizer::run,
    optimizer/optimizer.h:428-467: gradient of the 16 shots, direction, line
    search of forward-only trials, model and kernel export) through the
    command-line driver on a configuration directory written by
    problems.write_config, timed by the driver's own clock, which excludes
    setup and observed-data generation.  It inverts lambda only: the model's
    water layer has mu = 0, and the reference updates mu without a bound, so
    the first trial of a mu inversion makes mu negative there and the
    wavefield diverges (as it would in the reference)."""
    import re
    import subprocess
    import tempfile

    P = inv.problems.workload("c1", sigma=a.sigma)[0]
    exe = ROOT / "inv.cu_amd" / "inv.out"
    num = r"(-?(?:nan|inf|[0-9.eE+-]+))"
    with tempfile.TemporaryDirectory(prefix="inv_c1_") as tmp:
        root = inv.problems.write_config(P.with_(inv_lambda=1, inv_mu=0, inv_rho=0),
                                         Path(tmp) / "c1", mode=0, optimizer=0,
                                         inv_iteration=iterations, batch=16)
        t0 = time.perf_counter()
        p = subprocess.run([str(exe), str(root)], capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
    if p.returncode != 0:
        log(f"# c1 NLCG run failed: {p.stderr[-500:]}")
        return None
    m = re.search(r"Elapsed seconds: ([0-9.eE+-]+)\s+evaluations: (\d+)", p.stdout)
    final = re.search(r"Final misfit: " + num, p.stdout)
    restarts = p.stdout.count("restarting NLCG")
    secs = float(m.group(1)) if m else None
    rec = {"iterations": iterations,
           "seconds": round(secs, 3) if m else None,
           "seconds_per_iteration": round(secs / iterations, 3) if m else None,
           "forward_evaluations": int(m.group(2)) if m else None,
           "misfit_after": float(final.group(1)) if final else None,
           "restarts": restarts,
           "wall_s_including_setup": round(wall, 2),
           "note": "inv.out mode 0, NLCG (Polak-Ribiere), lambda inverted, 16 shots per "
                   f"launch, sigma {P.sigma}; misfit normalised to 1 at the start; an "
                   "evaluation is a forward of every shot (a gradient counts 2)"}
    log(f"# c1 NLCG run: {rec}")
    return rec


def full_gradient_c4(inv, a):
    """The 4096x2048 workload's whole gradient: Misfit::run(kernel=true)
    (misfit/misfit.h:15-51) over all 256 shots (forward, envelope misfit,
    adjoint, filter) through the command-line driver's mode 2 on a
    configuration directory written by problems.write_config.  Three shots
    per launch: their I16 snapshots (201 GB) and the observed data of all
    shots (42 GB) fit in HBM together (four shots' 268 GB would not), and a
    launch of three shots leaves less of its last round of tiles idle than
    one of two (on 24 of the shots: 32.9 s against 34.5 s).  Tiles are
    autotuned at the first model load.  The driver times the gradient itself (after setup and the
    observed-data forwards) and prints a progress line every 16 launches."""
    import re
    import subprocess
    import tempfile

    P = inv.problems.workload("c4", sigma=a.sigma)[0]
    exe = ROOT / "inv.cu_amd" / "inv.out"
    with tempfile.TemporaryDirectory(prefix="inv_c4_") as tmp:
        root = inv.problems.write_config(P, Path(tmp) / "c4", mode=2, batch=3, snap_format=3,
                                         autotune=1, progress=16)
        log(f"# c4 full gradient: {P.nsrc} shots through {exe.name}")
        t0 = time.perf_counter()
        p = subprocess.Popen([str(exe), str(root)], stdout=subprocess.DEVNULL,
                             stderr=subprocess.PIPE, text=True)
        err = []
        for line in p.stderr:
            err.append(line)
            log("#   " + line.rstrip())
        p.wait(timeout=60)
        wall = time.perf_counter() - t0
    m = re.search(r"gradient: ([0-9.eE+-]+) s, misfit (\S+), (\d+) shots", "".join(err))
    if p.returncode != 0 or not m:
        log(f"# c4 full gradient failed: rc {p.returncode}")
        return None
    secs, shots = float(m.group(1)), int(m.group(3))
    cells = shots * P.nx * P.nz * P.nt * 2
    rec = {"shots": shots, "seconds": round(secs, 2), "gcell_updates_per_s": round(cells / secs / 1e9, 2),
           "gradients_per_hr": round(3600 / secs, 2), "misfit": float(m.group(2)),
           "wall_s_including_setup": round(wall, 1),
           "note": "inv.out mode 2 on the written 4096x2048 configuration: all 256 shots, "
                   "3 per launch, I16 snapshots, envelope misfit, sigma " + str(P.sigma)}
    log(f"# c4 full gradient: {rec}")
    return rec


def lbfgs_c4(inv, a, iterations):
    """The 4096x2048 workload's inversion: `iterations` L-BFGS iterations
    (Optimizer::run, optimizer/optimizer.h:428-467, with lbfgs.h:18-91's
    direction, 5 stored pairs, and the backtracking line search of
    forward-only trials over all 256 shots, or every (256 / n)-th with
    --lbfgs-shots n) through the command-line driver
    on a configuration directory written by problems.write_config, three
    shots per launch with I16 snapshots as in full_gradient_c4.  Lambda only
    is inverted, as for c1 (the water layer has mu = 0, which an unbounded
    mu update drives negative).  The driver's own clock (Elapsed seconds)
    excludes setup and the observed-data forwards."""
    import re
    import subprocess
    import tempfile

    P = inv.problems.workload("c4", sigma=a.sigma)[0]
    if a.lbfgs_shots < P.nsrc:
        P = inv.problems.shots(P, np.arange(0, P.nsrc, P.nsrc // a.lbfgs_shots)[:a.lbfgs_shots])
    exe = ROOT / "inv.cu_amd" / "inv.out"
    num = r"(-?(?:nan|inf|[0-9.eE+-]+))"
    with tempfile.TemporaryDirectory(prefix="inv_c4_") as tmp:
        root = inv.problems.write_config(P.with_(inv_lambda=1, inv_mu=0, inv_rho=0),
                                         Path(tmp) / "c4", mode=0, optimizer=1, lbfgs_mem=5,
                                         inv_iteration=iterations, batch=3, snap_format=3,
                                         autotune=1, progress=16)
        log(f"# c4 L-BFGS: {iterations} iteration(s) over {P.nsrc} shots through {exe.name}")
        t0 = time.perf_counter()
        p = subprocess.Popen([str(exe), str(root)], stdout=subprocess.PIPE,
                             stderr=subprocess.STDOUT, text=True)
        lines = []
        for line in p.stdout:
            lines.append(line)
            log("#   " + line.rstrip())
        p.wait(timeout=60)
        wall = time.perf_counter() - t0
    out = "".join(lines)
    m = re.search(r"Elapsed seconds: ([0-9.eE+-]+)\s+evaluations: (\d+)", out)
    final = re.search(r"Final misfit: " + num, out)
    if p.returncode != 0 or not m:
        log(f"# c4 L-BFGS run failed: rc {p.returncode}")
        return None
    secs = float(m.group(1))
    its = [float(x) for x in re.findall(r"iteration \d+: ([0-9.eE+-]+) s since the start", out)]
    rec = {"iterations": iterations, "shots": P.nsrc, "seconds": round(secs, 1),
           "seconds_at_iteration_end": [round(x, 1) for x in its],
           "seconds_per_iteration": round(secs / iterations, 1),
           "forward_evaluations": int(m.group(2)),
           "misfit_after": float(final.group(1)) if final else None,
           "wall_s_including_setup": round(wall, 1),
           "note": f"inv.out mode 0, L-BFGS (5 pairs), lambda inverted, {P.nsrc} of the 256 "
                   f"shots, 3 per launch, I16 snapshots, envelope misfit, sigma {P.sigma}; "
                   "misfit normalised "
                   "to 1 at the start; an evaluation is a forward of every shot (a gradient "
                   "counts 2)"}
    log(f"# c4 L-BFGS run: {rec}")
    return rec


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): sharding the "
              "shots over several GPUs is designed but not built.", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as graft

    inv = graft.load_package()
    if a.serving_sweep:
        for s in inv.serving.sweep(a.serving_sweep, log=log,
                                   snap_format=inv.invhip.INV_SNAP_I16):
            print(json.dumps(s), flush=True)
        return
    if a.steps == 0 and (a.full_gradient or a.lbfgs):
        # only the whole-workload legs, one JSON line of their records
        rec = {}
        if a.full_gradient:
            rec["full_gradient_c4"] = full_gradient_c4(inv, a)
        if a.lbfgs:
            rec["lbfgs_c4"] = lbfgs_c4(inv, a, a.lbfgs)
        print(json.dumps(rec))
        return
    head, P, gpu_grad = run_workload(inv, a, a.workload, a.steps, a.warmup, headline=True)
    if head["roofline"]:
        # the achieved bandwidth against what a plain device copy reaches
        copy = copy_bandwidth(inv)
        head["roofline"]["copy_gbps"] = round(copy, 1)
        head["roofline"]["frac_of_copy"] = round(head["roofline"]["achieved"] / copy, 4)
        log(f"# device copy {copy:.1f} GB/s")
    extras_names = a.extras if a.extras is not None else ("c1,c2" if a.workload == "c4" else "")
    extras = {}
    for w in [x for x in extras_names.split(",") if x and x != a.workload]:
        rec, _, _ = run_workload(inv, a, w, 2, 1, headline=False)
        extras[w] = rec
    plain = opt = None
    if not a.no_plain and (a.workload == "c1" or "c1" in extras):
        plain = plain_c1(inv, a)
        opt = optimizer_c1(inv, a)

    cpu = parity = None
    if not a.no_cpu:
        info = cpu_info()
        # c0 in full; c1 on a sample of its shots; c2 / c4: short samples
        # (10-30 s of CPU work), plus one thread on a smaller one
        if a.workload == "c0":
            cpu, (fo, *ko) = cpu_baseline(P, P.nsrc, obs=gpu_grad[2])
            fg, kg, _ = gpu_grad
            parity = {"reference": "oracle FP32 gradient of all shots on the same observed data",
                      "math": ["exact", "fast"][a.math],
                      "misfit_rel_err": abs(fg - fo) / abs(fo)}
            for n, x, y in zip(("k_lambda", "k_mu", "k_rho"), kg, ko):
                parity[n + "_rel_l2"] = float(np.linalg.norm(x - y) / np.linalg.norm(y))
            log(f"# parity {parity}")
            one = cpu_baseline(P, 1, nt=200, sigma=0, threads=1)[0]
        elif a.workload == "c1":
            cpu = cpu_baseline(P, min(a.cpu_shots, P.nsrc))[0]
            one = cpu_baseline(P, 1, nt=400, threads=1)[0]
        elif a.workload == "c2":
            cpu = cpu_baseline(P, 8, nt=2000)[0]
            one = cpu_baseline(P, 1, nt=200, threads=1)[0]
        else:
            # the filter runs once per step of 4 x 10000 steps: left out of the
            # short samples, where it would dominate
            cpu = cpu_baseline(P, 1, nt=400, sigma=0)[0]
            one = cpu_baseline(P, 1, nt=40, sigma=0, threads=1)[0]
        cpu.update(info)
        cpu["one_thread"] = {k: one[k] for k in ("value", "cores", "sample")}
        log(f"# cpu baseline {cpu}")
    acc = None
    if not a.no_accuracy:
        acc = {w: accuracy(inv, w) for w in ("c4", "c2")
               if w == a.workload or (w in extras and a.workload == "c4")}
    full = full_gradient_c4(inv, a) if a.full_gradient and a.workload == "c4" else None
    lbfgs = lbfgs_c4(inv, a, a.lbfgs) if a.lbfgs > 0 and a.workload == "c4" else None
    serving = None
    if a.serve_requests > 0 and a.workload in ("c1", "c4"):
        serving = serving_leg(inv, a.serve_requests, a.serve_rate_x, inv.invhip.INV_SNAP_I16)

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic (checkerboard model; observed data modelled on the true model)"
                 if a.workload == "c0" else "synthetic (seeded Marmousi-shaped model; observed "
                                            "data modelled on the true model)"),
        "config": head["config"],
    }
    for k in ("gradients_per_hr", "gradients_per_hr_256_shots", "gradients_per_hr_64_shots",
              "gradients_per_hr_16_shots", "note"):
        if k in head:
            out[k] = head[k]
    out.update({"roofline": head["roofline"], "cpu_baseline": cpu, "plain": plain,
                "serving": serving, "accuracy": acc, "extra_workloads": extras or None,
                "optimizer_c1": opt})
    if full:
        out["full_gradient_c4"] = full
    if lbfgs:
        out["lbfgs_c4"] = lbfgs
    if parity:
        out["parity"] = parity
    print(json.dumps(out))


if __name__ == "__main__":
    main()

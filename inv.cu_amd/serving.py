"""Mixed-size serving load for the request server (inv_server_* of the C ABI).

Requests come from problems.serving_request (grid drawn from {256x256,
512x512, 1024x512, 2048x1024}, nt uniform in [1000, 6000], SH or P-SV and
forward or forward+adjoint with probability 1/2 each, 1-8 shots, a seeded
random layered model each).  Models are uploaded first and stay resident;
then requests are submitted at Poisson arrival times and latency is
t_done - t_submit on the server's clock.  The same arrival sequence can be
run through the continuous-batching server or the plain loop (one request at
a time, one shot per launch, no graph).
"""
from __future__ import annotations

import threading
import time

import numpy as np

from . import invhip, problems


def make_requests(n, seed=3, **kw):
    return [problems.serving_request(i, seed=seed, **kw) for i in range(n)]


def _submit(srv, req, mid):
    ncomp = 1 if req["sh"] else 2
    obs = None
    if req["adjoint"]:
        obs = np.zeros((len(req["src_ix"]), ncomp, len(req["rec_ix"]), req["nt"]), np.float32)
    return srv.submit(model=mid, nx=req["nx"], nz=req["nz"], dx=req["dx"], dt=req["dt"],
                      nt=req["nt"], src_ix=req["src_ix"], src_iz=req["src_iz"],
                      stf_x=req["stf_x"], stf_y=req["stf_y"], stf_z=req["stf_z"],
                      rec_ix=req["rec_ix"], rec_iz=req["rec_iz"], sh=req["sh"],
                      psv=req["psv"], adjoint=req["adjoint"], obs=obs, sfe=req["sfe"])


def run_load(reqs, rate=None, seed=3, plain=False, log=None, **server_kw):
    """Run `reqs` through a fresh server.  rate: Poisson arrivals per second
    (None: all submitted at once).  Returns a summary dict."""
    srv = invhip.Server(plain=plain, **server_kw)
    try:
        mids = []
        for q in reqs:
            lam, mu, rho = problems.layered_random(q["nx"], q["nz"], q["model_seed"])
            mids.append(srv.model_add(lam, mu, rho, q["nx"], q["nz"]))
        rng = np.random.default_rng([seed, 17])
        gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
        arrive = np.cumsum(gaps)
        ids = [None] * len(reqs)
        infos = [None] * len(reqs)
        ready = threading.Semaphore(0)

        def waiter():
            for k in range(len(reqs)):
                ready.acquire()
                infos[k], _, _ = srv.wait(ids[k])
                if log and (k + 1) % 8 == 0:
                    log(f"#   {'plain' if plain else 'batched'}: {k + 1} of {len(reqs)} "
                        "requests done")

        th = threading.Thread(target=waiter)
        th.start()
        t0 = srv.stats()["now"]
        w0 = time.perf_counter()
        for k, q in enumerate(reqs):
            delay = arrive[k] - (time.perf_counter() - w0)
            if delay > 0:
                time.sleep(delay)
            ids[k] = _submit(srv, q, mids[k])
            ready.release()
        th.join()
        st = srv.stats()
    finally:
        srv.close()
    lat = np.array([i["t_done"] - i["t_submit"] for i in infos])
    qwait = np.array([i["t_start"] - i["t_submit"] for i in infos])
    t_first = min(i["t_submit"] for i in infos)
    t_last = max(i["t_done"] for i in infos)
    span = t_last - t_first
    work = sum(problems.request_cell_updates(q) for q in reqs)
    out = {
        "mode": "plain" if plain else "batched",
        "requests": len(reqs),
        "rate_per_s": rate,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "mean_queue_s": float(qwait.mean()),
        "makespan_s": span,
        "requests_per_s": len(reqs) / span,
        "gcell_per_s": work / span / 1e9,
        "batches": st["batches"],
        "mean_requests_per_batch": float(np.mean([i["batch_requests"] for i in infos])),
        "peak_bytes": st["bytes_peak"],
        "budget_bytes": st["budget"],
        "submit_clock_offset_s": t0,
    }
    if log:
        log(f"# serving {out['mode']} rate={rate}: p50 {out['p50_latency_s']:.3f} s, "
            f"p99 {out['p99_latency_s']:.3f} s, {out['requests_per_s']:.2f} req/s, "
            f"{out['gcell_per_s']:.1f} Gcell/s, {out['batches']} batches")
    return out, infos


def sweep(n=32, rates_x=(0.25, 0.5, 1.0, 2.0, 4.0), seed=3, log=None, **server_kw):
    """C3 sweep: the plain loop's capacity from a back-to-back run of the n
    requests, then Poisson arrivals at rates_x times that capacity through the
    plain loop and the batched server.  Returns a list of summaries."""
    reqs = make_requests(n, seed=seed)
    base, _ = run_load(reqs, rate=None, seed=seed, plain=True, log=log, **server_kw)
    cap = base["requests_per_s"]
    out = [dict(base, rate_x=None, capacity_req_per_s=cap)]
    for x in rates_x:
        for plain in (True, False):
            s, _ = run_load(reqs, rate=x * cap, seed=seed, plain=plain, log=log, **server_kw)
            out.append(dict(s, rate_x=x, capacity_req_per_s=cap))
    return out

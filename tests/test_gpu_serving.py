"""Request server (inv_server_*): continuous batching of shot-solve requests.

* a request's results equal the same shots run through a plain context
  (forward traces bit for bit; adjoint kernels and misfit as the oracle's
  gradient within the FP32 tolerance);
* every request's results are the same alone and in a mix of concurrent,
  differently sized requests sharing launches (bit for bit: a slot's
  arithmetic does not depend on the batch it runs in);
* the plain one-request-at-a-time loop gives the same traces bit for bit and
  the same kernels to FP32 summation-order tolerance;
* admission keeps reservations within the device-memory budget, and the
  buffers the lanes really hold stay within their reservations;
* invalid requests are rejected at submit.
"""
import threading

import numpy as np
import pytest

from conftest import inv, oracle_problem, rel

pytestmark = pytest.mark.gpu

P = inv.problems
H = inv.invhip


def make_req(rid, nx, nz, nt, sh, adjoint, nshots, dx=10.0, dt=1e-3, sfe=10):
    """A small request with a seeded layered model; sources spread over x at
    3 cells depth, receivers every 3 cells on row 2."""
    xmax = (nx - 1) * dx
    src_ix = np.array([H.calc_idx((s + 0.5) * xmax / nshots, xmax, nx) for s in range(nshots)],
                      np.int32)
    stf = [H.source_ricker(nt, dt, 25.0, 0.04, 0.3 * s, 1e9) for s in range(nshots)]
    rec_ix = np.arange(2, nx - 2, 3, dtype=np.int32)
    return dict(rid=rid, model_seed=500 + rid, nx=nx, nz=nz, dx=dx, dt=dt, nt=nt, sh=int(sh),
                psv=int(not sh), adjoint=adjoint, sfe=sfe, src_ix=src_ix,
                src_iz=np.full(nshots, 3, np.int32),
                stf_x=np.stack([s[0] for s in stf]), stf_y=np.stack([s[1] for s in stf]),
                stf_z=np.stack([s[2] for s in stf]), rec_ix=rec_ix,
                rec_iz=np.full(len(rec_ix), 2, np.int32))


def obs_of(q):
    """Deterministic nonzero observed traces for an adjoint request."""
    rng = np.random.default_rng(q["rid"])
    ncomp = 1 if q["sh"] else 2
    return (1e-3 * rng.standard_normal((len(q["src_ix"]), ncomp, len(q["rec_ix"]), q["nt"]))
            ).astype(np.float32)


def submit(srv, q, mid):
    return srv.submit(model=mid, nx=q["nx"], nz=q["nz"], dx=q["dx"], dt=q["dt"], nt=q["nt"],
                      src_ix=q["src_ix"], src_iz=q["src_iz"], stf_x=q["stf_x"],
                      stf_y=q["stf_y"], stf_z=q["stf_z"], rec_ix=q["rec_ix"],
                      rec_iz=q["rec_iz"], sh=q["sh"], psv=q["psv"], adjoint=q["adjoint"],
                      obs=obs_of(q) if q["adjoint"] else None, sfe=q["sfe"])


def add_models(srv, reqs):
    out = []
    for q in reqs:
        lam, mu, rho = P.layered_random(q["nx"], q["nz"], q["model_seed"])
        out.append(srv.model_add(lam, mu, rho, q["nx"], q["nz"]))
    return out


def run_all(srv, reqs, together=True):
    """Results {rid: (info, traces, kernels)}: submitted all at once, or one
    at a time (each waited for before the next)."""
    mids = add_models(srv, reqs)
    res = {}
    if together:
        ids = [submit(srv, q, m) for q, m in zip(reqs, mids)]
        for q, i in zip(reqs, ids):
            res[q["rid"]] = srv.wait(i, timeout=120)
    else:
        for q, m in zip(reqs, mids):
            res[q["rid"]] = srv.wait(submit(srv, q, m), timeout=120)
    return res


MIX = [
    make_req(0, 64, 48, 300, sh=False, adjoint=True, nshots=2),
    make_req(1, 64, 48, 200, sh=False, adjoint=True, nshots=1),
    make_req(2, 64, 48, 260, sh=False, adjoint=True, nshots=3),
    make_req(3, 64, 48, 240, sh=True, adjoint=True, nshots=2),
    make_req(4, 64, 48, 300, sh=True, adjoint=True, nshots=1),
    make_req(5, 80, 40, 250, sh=False, adjoint=False, nshots=2),
    make_req(6, 80, 40, 170, sh=False, adjoint=False, nshots=3),
    make_req(7, 64, 48, 220, sh=True, adjoint=False, nshots=2),
]


def test_forward_traces_equal_a_plain_context(gpu):
    q = make_req(10, 64, 48, 230, sh=False, adjoint=False, nshots=3)
    lam, mu, rho = P.layered_random(q["nx"], q["nz"], q["model_seed"])
    with H.Server(max_lanes=1, max_slots=8) as srv:
        mid = srv.model_add(lam, mu, rho, q["nx"], q["nz"])
        info, tr, _ = srv.wait(submit(srv, q, mid), timeout=60)
    assert info["status"] == H.INV_REQ_DONE
    ctx = H.Context(0)
    ctx.grid_set(nx=q["nx"], nz=q["nz"], dx=q["dx"], dz=q["dx"], dt=q["dt"], nt=q["nt"],
                 sh=0, psv=1, adj=0, sfe=10, batch=3)
    ctx.model_upload(lam, mu, rho)
    ctx.sources_set(q["src_ix"], q["src_iz"], q["stf_x"], q["stf_y"], q["stf_z"])
    ctx.receivers_set(q["rec_ix"], q["rec_iz"])
    ctx.set_graphs(False)
    ctx.forward(0, 3, snapshots=False)
    for s in range(3):
        assert np.array_equal(tr[s], ctx.traces_get(s)), f"shot {s}"
    ctx.close()


@pytest.mark.parametrize("sh", [False, True])
def test_adjoint_request_matches_oracle(gpu, sh):
    q = make_req(20 + sh, 48, 40, 300, sh=sh, adjoint=True, nshots=2)
    lam, mu, rho = P.layered_random(q["nx"], q["nz"], q["model_seed"])
    obs = obs_of(q)
    with H.Server(max_lanes=1, max_slots=8, math=H.INV_MATH_EXACT) as srv:
        mid = srv.model_add(lam, mu, rho, q["nx"], q["nz"])
        rid = srv.submit(model=mid, nx=q["nx"], nz=q["nz"], dx=q["dx"], dt=q["dt"], nt=q["nt"],
                         src_ix=q["src_ix"], src_iz=q["src_iz"], stf_x=q["stf_x"],
                         stf_y=q["stf_y"], stf_z=q["stf_z"], rec_ix=q["rec_ix"],
                         rec_iz=q["rec_iz"], sh=q["sh"], psv=q["psv"], adjoint=True, obs=obs)
        info, tr, ks = srv.wait(rid, timeout=60)
    prob = P.Problem("req", q["nx"], q["nz"], np.float32(q["dx"]), np.float32(q["dx"]),
                     np.float32(q["dt"]), q["nt"], q["sh"], q["psv"], lam, mu, rho, lam, mu, rho,
                     q["src_ix"], q["src_iz"], q["stf_x"], q["stf_y"], q["stf_z"], q["rec_ix"],
                     q["rec_iz"])
    O = oracle_problem(prob, model="true", adj=1)
    f, kl, km, kr = O.gradient(obs, 0, 0)
    assert abs(info["misfit"] - f) <= 1e-5 * abs(f)
    names = ("k_lambda", "k_mu", "k_rho")
    for a, b, name in zip(ks, (kl, km, kr), names):
        if not np.any(b):
            continue
        assert rel(a, b) < 1e-4, f"{name}: {rel(a, b):.2e}"


def test_alone_equals_in_the_mix(gpu):
    with H.Server(max_lanes=3, max_slots=8, max_bypass=4) as srv:
        mixed = run_all(srv, MIX, together=True)
        st = srv.stats()
    with H.Server(max_lanes=1, max_slots=8) as srv:
        alone = run_all(srv, MIX, together=False)
    shared = 0
    for q in MIX:
        (ia, ta, ka), (ib, tb, kb) = mixed[q["rid"]], alone[q["rid"]]
        assert ia["status"] == ib["status"] == H.INV_REQ_DONE
        assert ib["batch_requests"] == 1
        shared += ia["batch_requests"] > 1
        assert np.array_equal(ta, tb), f"request {q['rid']}: traces differ"
        if q["adjoint"]:
            assert ia["misfit"] == ib["misfit"], f"request {q['rid']}: misfit differs"
            for x, y in zip(ka, kb):
                assert np.array_equal(x, y), f"request {q['rid']}: kernels differ"
            assert any(np.any(x) for x in ka)
    assert shared >= 2, "no launch was shared between requests"
    assert st["completed"] == len(MIX) and st["failed"] == 0
    assert st["bytes_actual"] <= st["bytes_held"]


def test_plain_loop_agrees_with_batched(gpu):
    reqs = [MIX[0], MIX[2], MIX[3], MIX[6]]
    with H.Server(max_lanes=2, max_slots=8) as srv:
        batched = run_all(srv, reqs, together=True)
    with H.Server(plain=True) as srv:
        plain = run_all(srv, reqs, together=True)
    for q in reqs:
        (ia, ta, ka), (ib, tb, kb) = batched[q["rid"]], plain[q["rid"]]
        assert ib["batch_shots"] == len(q["src_ix"]) and ib["batch_requests"] == 1
        assert np.array_equal(ta, tb)
        if q["adjoint"]:
            assert abs(ia["misfit"] - ib["misfit"]) <= 1e-12 * abs(ib["misfit"])
            for x, y in zip(ka, kb):
                assert rel(x, y) < 1e-6


def test_budget_bounds_reservations(gpu):
    # each batch reserves about 85 MiB (2 MiB allocator rounding per buffer
    # dominates at this size): a 128 MiB budget admits one batch at a time
    reqs = [make_req(30 + k, 96, 64, 400, sh=False, adjoint=True, nshots=2) for k in range(4)]
    with H.Server(max_lanes=4, max_slots=2, hbm_budget=128 << 20) as srv:
        res = run_all(srv, reqs, together=True)
        st = srv.stats()
    assert st["budget"] == 128 << 20
    assert st["completed"] == 4 and st["bytes_peak"] <= st["budget"]
    spans = sorted((r[0]["t_start"], r[0]["t_done"]) for r in res.values())
    assert all(r[0]["status"] == H.INV_REQ_DONE for r in res.values())
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1, "two batches ran at once beyond the budget"


def test_budget_rejects_oversized_request(gpu):
    q = make_req(40, 256, 256, 2000, sh=False, adjoint=True, nshots=4)
    with H.Server(max_lanes=1, max_slots=8, hbm_budget=64 << 20) as srv:
        lam, mu, rho = P.layered_random(q["nx"], q["nz"], q["model_seed"])
        mid = srv.model_add(lam, mu, rho, q["nx"], q["nz"])
        with pytest.raises(H.InvError) as e:
            submit(srv, q, mid)
        assert e.value.code == -4


def test_invalid_requests_rejected(gpu):
    q = make_req(41, 48, 40, 200, sh=False, adjoint=True, nshots=1)
    with H.Server(max_lanes=1, max_slots=2) as srv:
        lam, mu, rho = P.layered_random(q["nx"], q["nz"], q["model_seed"])
        mid = srv.model_add(lam, mu, rho, q["nx"], q["nz"])
        with pytest.raises(H.InvError):
            submit(srv, q, mid + 99)                     # unknown model
        with pytest.raises(H.InvError):                  # nt % sfe != 0 with the adjoint
            submit(srv, make_req(41, 48, 40, 205, sh=False, adjoint=True, nshots=1), mid)
        with pytest.raises(H.InvError):
            submit(srv, dict(q, nx=50), mid)             # grid differs from the model's
        big = make_req(42, 48, 40, 200, sh=False, adjoint=False, nshots=3)
        with pytest.raises(H.InvError):
            submit(srv, big, mid)                        # more shots than slots
        rid = submit(srv, q, mid)
        info, _, _ = srv.wait(rid, timeout=60)
        assert info["status"] == H.INV_REQ_DONE
        srv.model_remove(mid)
        with pytest.raises(H.InvError):
            srv.model_remove(mid)


def test_concurrent_submitters(gpu):
    reqs = MIX[:6]
    with H.Server(max_lanes=2, max_slots=8) as srv:
        mids = add_models(srv, reqs)
        out = {}

        def client(k):
            q = reqs[k]
            out[q["rid"]] = srv.wait(submit(srv, q, mids[k]), timeout=120)

        th = [threading.Thread(target=client, args=(k,)) for k in range(len(reqs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
    with H.Server(max_lanes=1, max_slots=8) as srv:
        alone = run_all(srv, reqs, together=False)
    for q in reqs:
        assert np.array_equal(out[q["rid"]][1], alone[q["rid"]][1])

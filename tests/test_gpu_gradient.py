"""GPU parity of the whole gradient evaluation (Misfit::run(true),
misfit/misfit.h:15-51): forward, waveform misfit, adjoint, Gaussian filter."""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel

pytestmark = pytest.mark.gpu


def _obs_from_true(P):
    """Observed data: forward on the true model (Misfit::init, misfit.h:135-148)."""
    O = oracle_problem(P, model="true")
    obs = []
    for isrc in range(P.nsrc):
        tr, _, _ = O.forward(isrc)
        obs.append(oracle_traces(tr, P))
    return np.stack(obs).astype(np.float32)


@pytest.mark.parametrize("math", [0, 1], ids=["exact", "fast"])
@pytest.mark.parametrize("variant", [0, 1, (8, 2)], ids=["fused", "split", "wide"])
@pytest.mark.parametrize("batch", [1, 3])
@pytest.mark.parametrize("sigma", [0, 3])
def test_gradient_matches_oracle(gpu, batch, sigma, variant, math):
    P = inv.problems.tiny(sh=1, psv=1, nsrc=3)
    obs = _obs_from_true(P)
    print(f"gradient batch={batch} sigma={sigma} variant={variant} math={math}", flush=True)
    ctx = P.context(batch=batch)
    if isinstance(variant, tuple):
        ctx.set_tile(*variant)
    else:
        ctx.set_variant(variant)
    ctx.set_math(math)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=0, sigma=sigma)
    kl, km, kr = ctx.kernels_get()
    O = oracle_problem(P)
    f32, okl, okm, okr = O.gradient(obs, 0, sigma, dtype=np.float32)
    f64, dkl, dkm, dkr = O.gradient(obs.astype(np.float64), 0, sigma, dtype=np.float64)
    assert abs(f - f64) / abs(f64) < 1e-5
    for a, b, c, name in ((kl, okl, dkl, "k_lambda"), (km, okm, dkm, "k_mu"),
                          (kr, okr, dkr, "k_rho")):
        e32, e64 = rel(a, b), rel(a, c)
        print(f"  {name}: vs f32 oracle {e32:.2e}, vs f64 oracle {e64:.2e}")
        if batch == 1 and math == 0:
            assert np.array_equal(a, b), f"{name} plain path not bit-exact ({e32:.3e})"
        assert e32 < 1e-4 and e64 < 1e-4
    ctx.close()


def test_misfit_waveform_matches_oracle(gpu):
    P = inv.problems.tiny(sh=1, psv=1, nsrc=2)
    obs = _obs_from_true(P)
    ctx = P.context(batch=2)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    ctx.forward(0, 2, snapshots=False)
    tot, chi = ctx.misfit(0, 2, kind=0)
    import pyoracle
    ref_tot = 0.0
    for s in range(2):
        syn = ctx.traces_get(s)
        ad = ctx.adstf_get(s)
        for c in range(P.ncomp):
            for r in range(P.nrec):
                x, a = pyoracle.waveform(syn[c, r], obs[s, c, r], P.dt)
                assert np.array_equal(a, ad[c, r])
                assert chi[s, r, c] == np.float64(np.float32(x))
                ref_tot += x
    assert abs(tot - ref_tot) <= 1e-12 * abs(ref_tot)
    ctx.close()


def test_filter_matches_oracle(gpu):
    import pyoracle
    rng = np.random.default_rng(3)
    nx, nz = 70, 53
    a = rng.standard_normal(nx * nz).astype(np.float32)
    ctx = inv.invhip.Context(0)
    dev = ctx.alloc(a.nbytes)
    for sigma in (1, 5, 20):
        ctx.h2d(dev, a)
        ctx.filter_gauss(dev, nx, nz, sigma)
        out = ctx.d2h(dev, a.shape)
        ref = pyoracle.gauss(a, nx, nz, sigma)
        assert np.array_equal(out, ref), f"sigma {sigma}: rel {rel(out, ref):.3e}"
    ctx.close()


def test_vector_ops(gpu):
    rng = np.random.default_rng(4)
    n = 100003
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    ctx = inv.invhip.Context(0)
    da, db, dc = (ctx.alloc(4 * n) for _ in range(3))
    ctx.h2d(da, a)
    ctx.h2d(db, b)
    ctx.vec_axpby(dc, 0.5, da, -2.0, db, n)
    assert np.array_equal(ctx.d2h(dc, n), np.float32(0.5) * a + np.float32(-2.0) * b)
    ctx.vec_scale(dc, 3.0, da, n)
    assert np.array_equal(ctx.d2h(dc, n), np.float32(3.0) * a)
    ctx.vec_mul(dc, da, db, n)
    assert np.array_equal(ctx.d2h(dc, n), a * b)
    ctx.vec_copy(dc, db, n)
    assert np.array_equal(ctx.d2h(dc, n), b)
    d = ctx.vec_dot(da, db, n)
    ref = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    assert abs(d - ref) <= 1e-12 * np.sqrt(n) * np.linalg.norm(a) * np.linalg.norm(b)
    assert ctx.vec_amax(da, n) == np.abs(a).max()
    ctx.close()


@pytest.mark.parametrize("nt", [300, 320, 375, 2000, 311, 1130])
def test_misfit_envelope_matches_oracle(gpu, nt):
    """Envelope adjoint sources and misfits against the oracle (FP64 Hilbert
    transform): the GPU transform is FP32, so agreement is to tolerance.
    311 (prime) and 1130 (2 5 113) go through Bluestein transforms of length
    1024 and 4096."""
    import pyoracle
    P = inv.problems.tiny(sh=1, psv=1, nsrc=2, nt=nt, sfe=5 if nt % 5 == 0 else 1)
    obs = _obs_from_true(P)
    ctx = P.context(batch=2)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    ctx.forward(0, 2, snapshots=False)
    tot, chi = ctx.misfit(0, 2, kind=1)
    ref_tot = 0.0
    worst = 0.0
    for s in range(2):
        syn = ctx.traces_get(s)
        ad = ctx.adstf_get(s)
        for c in range(P.ncomp):
            for r in range(P.nrec):
                x, a = pyoracle.envelope(syn[c, r].astype(np.float64),
                                         obs[s, c, r].astype(np.float64), P.dt, np.float64)
                worst = max(worst, rel(ad[c, r], a))
                assert abs(chi[s, r, c] - x) <= 1e-4 * abs(x) + 1e-12
                ref_tot += x
    print(f"nt={nt}: worst adstf rel err vs FP64 oracle {worst:.2e}")
    assert worst < 1e-4
    assert abs(tot - ref_tot) <= 1e-5 * abs(ref_tot)
    ctx.close()


def test_envelope_rejects_unsupported_length(gpu):
    """Traces longer than 2^20 samples are refused before anything runs."""
    nt = (1 << 20) + 1
    P = inv.problems.tiny(sh=1, psv=0, nsrc=1, nx=24, nz=20, nrec=2, nt=nt, sfe=1)
    ctx = P.context(batch=1, adj=0)
    d = [ctx.alloc(4 * nt) for _ in range(3)]
    with pytest.raises(inv.invhip.InvError) as e:
        ctx.misfit_trace(1, d[0], d[1], d[2])
    assert e.value.code == -5
    ctx.close()


def _envelope_pair(nt, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(nt) * 1e-3
    syn = (np.sin(40 * t) * np.exp(-3 * t) + 0.1 * rng.standard_normal(nt)).astype(np.float32)
    obs = (np.sin(41 * t) * np.exp(-3 * t) + 0.1 * rng.standard_normal(nt)).astype(np.float32)
    return syn, obs


@pytest.mark.parametrize("nt", [8190, 4099, 8320, 16640, 20000])
def test_envelope_long_bluestein_trace(gpu, nt):
    """Long traces on one device trace against the FP64 oracle: the longest
    Bluestein transform held in LDS (8190 = 2 3^2 5 7 13: length 16384), a
    prime length, and transforms run through global memory: Bluestein of
    length 32768 (8320 = 2^7 5 13) and 65536 (16640 = 2^8 5 13), and the
    direct mixed-radix transform of 20000 = 2^5 5^4."""
    import pyoracle
    P = inv.problems.tiny(sh=1, psv=0, nsrc=1, nx=24, nz=20, nrec=2, nt=nt, sfe=1)
    ctx = P.context(batch=1, adj=0)
    syn, obs = _envelope_pair(nt, nt)
    d = [ctx.alloc(4 * nt) for _ in range(3)]
    ctx.h2d(d[0], syn)
    ctx.h2d(d[1], obs)
    print(f"envelope trace nt={nt}", flush=True)
    chi = ctx.misfit_trace(1, d[0], d[1], d[2])
    x, a = pyoracle.envelope(syn.astype(np.float64), obs.astype(np.float64), P.dt, np.float64)
    e = rel(ctx.d2h(d[2], nt), a)
    print(f"nt={nt}: envelope adjoint source rel err vs FP64 {e:.2e}, misfit {abs(chi - x) / x:.2e}")
    assert abs(chi - x) <= 1e-4 * x
    assert e < 1e-4
    ctx.close()


def test_envelope_global_transforms_in_chunks(gpu):
    """The envelope misfit of 960 traces of 8320 samples (Bluestein of length
    32768 through global memory, 860 traces per launch: two launches): the
    misfit of every trace, and the adjoint sources of the first and last
    trace of each launch and of both shots, against the FP64 oracle; the
    total is the sum of the traces'."""
    import pyoracle
    nt = 8320
    P = inv.problems.tiny(sh=0, psv=1, nsrc=2, nx=250, nz=20, nrec=240, nt=nt, sfe=10)
    tctx = P.context(batch=2, model="true", adj=0)
    tctx.forward(0, 2, snapshots=False)
    obs = np.stack([tctx.traces_get(b) for b in range(2)])
    tctx.close()
    ctx = P.context(batch=2)
    for s in range(2):
        ctx.obs_set(s, obs[s])
    ctx.forward(0, 2, snapshots=False)
    print("envelope misfit of 960 traces, nt 8320", flush=True)
    tot, chi = ctx.misfit(0, 2, kind=1)
    syn = [ctx.traces_get(s) for s in range(2)]
    ad = [ctx.adstf_get(s) for s in range(2)]
    ctx.close()
    assert abs(tot - chi.sum()) <= 1e-12 * abs(tot)
    flat = [(s, c, r) for s in range(2) for c in range(2) for r in range(240)]
    check = {0, 1, 859, 860, 861, 959, 480, 479}
    worst = 0.0
    for t, (s, c, r) in enumerate(flat):
        if t not in check and t % 37:
            continue
        x, a = pyoracle.envelope(syn[s][c, r].astype(np.float64),
                                 obs[s, c, r].astype(np.float64), P.dt, np.float64)
        assert abs(chi[s, r, c] - x) <= 1e-4 * abs(x) + 1e-12, f"trace {t}"
        if t in check:
            worst = max(worst, rel(ad[s][c, r], a))
    print(f"worst adjoint source rel err vs FP64 {worst:.2e}")
    assert worst < 1e-4


def test_incremental_run_equals_whole_run(gpu):
    """inv_forward_begin + inv_run_advance in chunks gives the same bits as one
    inv_forward; the state between chunks is the oracle's state at that step."""
    P = inv.problems.tiny(sh=1, psv=1, nsrc=2)
    O = oracle_problem(P)
    a = P.context(batch=1)
    a.set_math(0)
    a.forward(1, 1, snapshots=True)
    b = P.context(batch=1)
    b.set_math(0)
    b.forward_begin(1, 1, snapshots=True)
    pos = 0
    for chunk in (7, 50, 100, 1000):
        pos = b.advance(chunk)
        if pos < P.nt:
            _, _, st = O.forward(1, nsteps=pos)
            assert np.array_equal(b.state_get(0), st)
    assert pos == P.nt
    assert np.array_equal(a.traces_get(0), b.traces_get(0))
    assert np.array_equal(a.snapshots_get(0), b.snapshots_get(0))
    assert np.array_equal(a.state_get(0), b.state_get(0))
    a.close()
    b.close()


def test_misfit_trace_on_device_pointers(gpu):
    import pyoracle
    P = inv.problems.tiny(sh=1, psv=0, nsrc=1, nt=400)
    ctx = P.context(batch=1)
    rng = np.random.default_rng(9)
    syn = rng.standard_normal(P.nt).astype(np.float32)
    obs = rng.standard_normal(P.nt).astype(np.float32)
    d = [ctx.alloc(4 * P.nt) for _ in range(3)]
    ctx.h2d(d[0], syn)
    ctx.h2d(d[1], obs)
    chi = ctx.misfit_trace(0, d[0], d[1], d[2])
    x, a = pyoracle.waveform(syn, obs, P.dt)
    assert chi == np.float64(np.float32(x))
    assert np.array_equal(ctx.d2h(d[2], P.nt), a)
    chi = ctx.misfit_trace(1, d[0], d[1], d[2])
    x, a = pyoracle.envelope(syn.astype(np.float64), obs.astype(np.float64), P.dt, np.float64)
    assert abs(chi - x) <= 1e-4 * x
    assert rel(ctx.d2h(d[2], P.nt), a) < 1e-4
    ctx.close()


def test_model_swap_after_async_forward(gpu):
    """Observed data modelled on the device, then the initial model uploaded
    right away (the bench's sequence): the upload must wait for the forward
    still running on the context stream.  The misfit must equal the oracle's."""
    P = inv.problems.marmousi(nx=200, nz=80, nt=600, nsrc=4, nrec=50, rec_step=4, f0=8.0,
                              t0=0.15)
    ctx = P.context(batch=4, model="true")
    ctx.forward(0, 4, snapshots=False)
    ctx.obs_from_traces(0, 4)
    P.load_model(ctx, "init")
    f1 = ctx.gradient(0, 0, kernel=False)
    f2 = ctx.gradient(0, 0, kernel=False)
    assert f1 == f2
    O = oracle_problem(P, model="true")
    obs = np.stack([oracle_traces(O.forward(s)[0], P) for s in range(P.nsrc)])
    f64, _, _, _ = oracle_problem(P).gradient(obs.astype(np.float64), 0, 0, kernel=False,
                                             dtype=np.float64)
    assert abs(f1 - f64) <= 1e-5 * abs(f64)
    ctx.close()


@pytest.mark.parametrize("variant", [0, 1, (4, 4)], ids=["fused", "split", "wide"])
def test_graph_replay_equals_direct(gpu, variant):
    """Captured-and-replayed runs give the same bits as direct launches, over
    several gradients and alternating shot batches (graph per batch)."""
    P = inv.problems.tiny(sh=1, psv=1, nsrc=4)
    obs = _obs_from_true(P)
    res = []
    for graphs in (False, True):
        ctx = P.context(batch=2)
        if isinstance(variant, tuple):
            ctx.set_tile(*variant)
        else:
            ctx.set_variant(variant)
        ctx.set_graphs(graphs)
        for s in range(P.nsrc):
            ctx.obs_set(s, obs[s])
        out = []
        for _ in range(3):
            f = ctx.gradient(0, 2)
            out.append((f, ctx.kernels_get(), ctx.state_get(1)))
        res.append(out)
        ctx.close()
    for (fa, ka, sa), (fb, kb, sb) in zip(*res):
        assert fa == fb
        for x, y in zip(ka, kb):
            assert np.array_equal(x, y)
        assert np.array_equal(sa, sb)


@pytest.mark.parametrize("fmt", [0, 1, 2, 3], ids=["f32", "f16", "bf16", "i16"])
@pytest.mark.parametrize("nz", [40, 41], ids=["even-nz", "odd-nz"])
def test_kernel_accumulation_matches_oracle(gpu, nz, fmt):
    """Adjoint kernels against the oracle's adjoint driven with the GPU's own
    (decoded) snapshots: equal value for value in exact math, for every
    snapshot format.  Even nz runs the column-strip accumulation kernel
    (k_interact_col), odd nz the per-cell one (k_interact)."""
    P = inv.problems.tiny(nz=nz, sh=1, psv=1, nsrc=1)
    print(f"accumulation nz={nz} fmt={fmt}", flush=True)
    ctx = P.context(batch=1)
    ctx.set_math(0)
    ctx.set_snapshots(fmt, 0)
    ctx.kernels_zero()
    ctx.forward(0, 1, snapshots=True)
    tr = ctx.traces_get(0)
    adstf = np.ascontiguousarray(tr[:, :, ::-1] * np.float32(0.5))
    ctx.adstf_set(0, adstf)
    ctx.adjoint(0, 1)
    sn = ctx.snapshots_get(0)
    kl, km, kr = ctx.kernels_get()
    ctx.close()
    O = oracle_problem(P)
    ks = [np.zeros(P.N, np.float32) for _ in range(3)]
    comps = ["y", "x", "z"]
    O.adjoint({c: adstf[i] for i, c in enumerate(comps)}, sn, kernels=ks)
    for a, b, name in ((kl, ks[0], "k_lambda"), (km, ks[1], "k_mu"), (kr, ks[2], "k_rho")):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b), f"{name}: rel L2 {rel(a, b):.3e}"

"""The persistent two-step kernel (inv_set_persistent): one launch runs every
pair of steps of a forward, or of an adjoint between kernel accumulations,
with each tile's wavefields held on the chip and only halo rows exchanged
between neighbouring tiles.  No counterpart in the reference, which launches
every half step (solver/fdm.h:534-551).  Bars:

  * exact math: traces, final wavefields and snapshots equal the FP32
    oracle's value for value (SH, P-SV, both; point forces and explosions;
    sources and receivers on tile row edges, in the first and last columns,
    a last tile of fewer than 8 rows);
  * any math mode: gradients and wavefields equal those of the two-step
    kernel launched per pair (8 x 4 tile) bit for bit, FP32 and I16
    snapshots, graphs and direct launches, several shots per launch, runs
    advanced in chunks, the 200x200 four-shot explosive workload;
  * configurations it does not run are rejected; a wait that gives up ends
    the run with an error instead of a hang, and the context runs again.
"""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel

pytestmark = pytest.mark.gpu

H = inv.invhip


def res_problem(sh=1, psv=1, nt=170, explosive=False, nx=70, nz=200):
    """Tiles of 8 rows (the last of 6), one tile across the 200 columns;
    sources and receivers on and next to tile row edges and grid columns
    0..3 and 196..199 (the lanes beside them read clamped columns)."""
    P = inv.problems.tiny(nx=nx, nz=nz, nt=nt, nsrc=4, nrec=12, sh=sh, psv=psv)
    P = P.with_(src_ix=np.array([15, 16, 40, 64], np.int32),
                src_iz=np.array([100, 3, 196, 52], np.int32),
                rec_ix=np.array([7, 8, 9, 23, 24, 31, 32, 47, 56, 63, 64, 69], np.int32),
                rec_iz=np.array([2, 100, 198, 5, 150, 0, 199, 60, 4, 197, 120, 80], np.int32))
    if explosive:
        P = P.with_(src_type=np.array([1, 0, 1, 1], np.int32))
    return P


def _forward_all(ctx, P, batch):
    out = []
    for first in range(0, P.nsrc, batch):
        ctx.forward(first, batch, snapshots=True)
        for s in range(batch):
            out.append((ctx.traces_get(s), ctx.state_get(s), ctx.snapshots_get(s)))
    ctx.sync()
    return out


@pytest.mark.parametrize("case", [dict(sh=1, psv=0), dict(sh=0, psv=1), dict(sh=1, psv=1),
                                  dict(sh=0, psv=1, explosive=True)],
                         ids=["sh", "psv", "sh+psv", "psv-explosive"])
def test_resident_forward_matches_oracle(gpu, case):
    P = res_problem(**case)
    print(f"resident forward {case}", flush=True)
    ctx = P.context(batch=2)
    ctx.set_math(H.INV_MATH_EXACT)
    ctx.set_pair(8, 4)
    ctx.set_persistent(1)
    ctx.set_graphs(False)  # launches inside a captured graph are not sampled
    ctx.prof_enable(1)
    got = _forward_all(ctx, P, 2)
    n, _ = ctx.prof_read(H.INV_PROF_RESIDENT)
    ctx.close()
    assert n == 2 * P.nt, "the persistent kernel ran every step"
    O = oracle_problem(P)
    for s in range(P.nsrc):
        otr, osn, ost = O.forward(s, snapshots=True)
        tr, st, sn = got[s]
        assert np.array_equal(tr, oracle_traces(otr, P)), f"traces of shot {s}"
        assert np.array_equal(st, ost), f"wavefields of shot {s}"
        assert np.array_equal(sn, osn), f"snapshots of shot {s}"


def _obs(P):
    O = oracle_problem(P, model="true")
    return np.stack([oracle_traces(O.forward(s)[0], P) for s in range(P.nsrc)]).astype(np.float32)


def _gradient(P, obs, batch, persist, fmt, graphs=True, math=H.INV_MATH_FAST, misfit=0, rows=8):
    ctx = P.context(batch=batch)
    ctx.set_math(math)
    ctx.set_graphs(graphs)
    ctx.set_snapshots(fmt, 0)
    ctx.set_pair(rows, 4)
    ctx.set_persistent(persist)
    ctx.prof_enable(1 if persist else 0)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=misfit, sigma=0)
    k = ctx.kernels_get()
    ctx.forward(0, batch, snapshots=False)
    tr = [ctx.traces_get(b) for b in range(batch)]
    st = [ctx.state_get(b) for b in range(batch)]
    ctx.sync()
    n = ctx.prof_read(H.INV_PROF_RESIDENT)[0]
    ctx.close()
    return f, k, tr, st, n


@pytest.mark.parametrize("waves", ["psv", "sh", "sh+psv"])
@pytest.mark.parametrize("rows", [8, 4])
@pytest.mark.parametrize("batch", [1, 3])
@pytest.mark.parametrize("fmt,graphs,math", [(H.INV_SNAP_F32, True, H.INV_MATH_FAST),
                                             (H.INV_SNAP_F32, False, H.INV_MATH_EXACT),
                                             (H.INV_SNAP_I16, True, H.INV_MATH_FAST),
                                             (H.INV_SNAP_F16, False, H.INV_MATH_FAST)],
                         ids=["f32-graphs-fast", "f32-direct-exact", "i16-graphs-fast",
                              "f16-direct-fast"])
def test_resident_gradient_equals_pair(gpu, waves, rows, batch, fmt, graphs, math):
    """Gradient (forward with snapshots, misfit, adjoint with kernel
    accumulations every sfe steps: in the launch for one wave type, after
    launches of the pairs up to each accumulation for both) and a following
    forward: the per-pair launches' results bit for bit (the same
    240-column snapshot segments)."""
    sh, psv = {"psv": (0, 1), "sh": (1, 0), "sh+psv": (1, 1)}[waves]
    P = res_problem(sh=sh, psv=psv, nt=300, explosive=bool(psv))
    P = inv.problems.shots(P, np.arange(batch))
    obs = _obs(P)
    print(f"resident gradient {waves} rows={rows} batch={batch} fmt={fmt} graphs={graphs} "
          f"math={math}", flush=True)
    f1, k1, t1, s1, n1 = _gradient(P, obs, batch, 1, fmt, graphs, math, rows=rows)
    f0, k0, t0, s0, _ = _gradient(P, obs, batch, 0, fmt, graphs, math, rows=rows)
    assert n1 > 0 or graphs, "persistent launches ran"
    assert f1 == f0
    assert np.abs(k0[1]).max() > 0 and np.abs(k0[2]).max() > 0  # k_mu, k_rho
    for a, b in zip(k1, k0):  # k_lambda is zero for SH alone
        assert np.array_equal(a, b), f"kernels differ, rel L2 {rel(a, b):.3e}"
    for a, b in zip(t1 + s1, t0 + s0):
        assert np.array_equal(a, b)


def test_resident_gradient_matches_oracle(gpu):
    """Exact math, one shot per launch: the FP32 oracle's gradient value for
    value."""
    P = res_problem(sh=1, psv=1, nt=200, explosive=True)
    P = inv.problems.shots(P, [0])
    obs = _obs(P)
    f, k, _, _, _ = _gradient(P, obs, 1, 1, H.INV_SNAP_F32, math=H.INV_MATH_EXACT)
    O = oracle_problem(P)
    f32, *k32 = O.gradient(obs, 0, 0, dtype=np.float32)
    assert abs(f - f32) / abs(f32) < 1e-6
    for a, b, name in zip(k, k32, ("k_lambda", "k_mu", "k_rho")):
        assert np.array_equal(a, b), f"{name} rel L2 {rel(a, b):.3e}"


@pytest.mark.parametrize("chunks", [[1, 40, 41, 88], [7, 7, 7, 149]], ids=["odd-first", "sevens"])
def test_resident_chunked_runs_equal_whole(gpu, chunks):
    """Forward and adjoint advanced in chunks (a persistent launch per chunk
    of pairs, odd steps alone) equal whole runs."""
    P = res_problem(sh=0, psv=1, nt=170, explosive=True)
    P = inv.problems.shots(P, [0, 1])
    print(f"resident chunks {chunks}", flush=True)

    def run(chunked):
        ctx = P.context(batch=2)
        ctx.set_pair(8, 4)
        ctx.set_persistent(1)
        if chunked:
            ctx.forward_begin(0, 2, snapshots=True)
            for n in chunks:
                ctx.advance(n)
        else:
            ctx.forward(0, 2, snapshots=True)
        tr = [ctx.traces_get(b) for b in range(2)]
        st = [ctx.state_get(b) for b in range(2)]
        ctx.adstf_set(0, np.ascontiguousarray(tr[0][:, :, ::-1]))
        ctx.adstf_set(1, np.ascontiguousarray(tr[1][:, :, ::-1]))
        ctx.kernels_zero()
        if chunked:
            ctx.adjoint_begin(0, 2)
            for n in chunks:
                ctx.advance(n)
        else:
            ctx.adjoint(0, 2)
        k = ctx.kernels_get()
        ctx.sync()
        ctx.close()
        return tr, st, k

    a, b = run(True), run(False)
    for x, y in zip(a[0] + a[1] + list(a[2]), b[0] + b[1] + list(b[2])):
        assert np.array_equal(x, y)


def test_resident_c0_workload_equals_pair(gpu):
    """The 200x200 checkerboard workload with four explosive shots in one
    launch (nt 2000, FP32 snapshots, fast math, graphs): the gradient of the
    persistent path equals the per-pair launches' bit for bit."""
    P, B, fmt, ckpt = inv.problems.workload("c0", sigma=0)
    ctx = P.context(batch=B, model="true")
    ctx.forward(0, B, snapshots=False)
    obs = np.stack([ctx.traces_get(b) for b in range(B)])
    ctx.close()
    print("resident c0", flush=True)
    f1, k1, _, s1, n1 = _gradient(P, obs, B, 1, fmt)
    f0, k0, _, s0, _ = _gradient(P, obs, B, 0, fmt)
    assert f1 == f0
    for a, b in zip(k1 + s1, k0 + s0):
        assert np.array_equal(a, b)


def test_resident_rejects_what_it_cannot_run(gpu):
    P = res_problem(sh=0, psv=1, nt=20, nz=256)
    ctx = P.context(batch=1)
    ctx.set_pair(8, 4)
    ctx.set_persistent(1)
    with pytest.raises(H.InvError, match="nz <= 240"):
        ctx.forward(0, 1, snapshots=False)
    ctx.close()
    P = res_problem(sh=0, psv=1, nt=20)
    ctx = P.context(batch=1)
    ctx.set_pair(8, 2)
    ctx.set_persistent(1)
    with pytest.raises(H.InvError, match="8 rows x 4 columns"):
        ctx.forward(0, 1, snapshots=False)
    ctx.set_pair(8, 4)
    ctx.set_snapshots(H.INV_SNAP_F16, 10)
    with pytest.raises(H.InvError, match="checkpoint"):
        ctx.forward(0, 1, snapshots=True)
    ctx.close()


def test_resident_wait_that_gives_up_is_an_error(gpu):
    """Mode 2 gives up every wait at once (the kernel ends early, every
    workgroup after the same barrier): the next synchronising call reports
    it, and the context then runs normally again."""
    P = res_problem(sh=0, psv=1, nt=40)
    ctx = P.context(batch=1)
    ctx.set_pair(8, 4)
    ctx.set_persistent(2)
    ctx.forward(0, 1, snapshots=False)
    with pytest.raises(H.InvError, match="gave up waiting"):
        ctx.sync()
    ctx.set_persistent(1)
    ctx.forward(0, 1, snapshots=False)
    tr = ctx.traces_get(0)
    ctx.sync()
    ctx.set_persistent(0)
    ctx.forward(0, 1, snapshots=False)
    tr0 = ctx.traces_get(0)
    ctx.close()
    assert np.array_equal(tr, tr0)

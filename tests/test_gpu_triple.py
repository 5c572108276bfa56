"""The three-step kernel (inv_set_triple): three time steps per launch with
the tile's stress and velocity kept in LDS between them.  No counterpart in
the reference, which launches every half step (solver/fdm.h:534-551).  Bars:

  * exact math: traces, final wavefields and snapshots equal the FP32
    oracle's value for value (SH, P-SV, both; point forces and explosions;
    sources and receivers on the three-step tiles' row and column edges),
    with and without a two-step tile for the steps that store snapshots;
  * any math mode: gradients (forward, misfit, adjoint with kernel
    accumulations) and a following forward equal those without three-step
    launches bit for bit with FP32 snapshots, graph replays and direct
    launches, whole and chunked runs, 1 and 3 shots per launch; 16-bit
    snapshots to their format's tolerance (a snapshot step may fall to a
    one-step launch, whose row segments are 120 columns against 116);
  * the 4096x2048 grid: wavefields and traces equal the two-step path's;
  * unsupported tiles and odd nz are rejected.
"""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel

pytestmark = pytest.mark.gpu

H = inv.invhip
TRIPS = [12, 8, 16, 20]


def edge_problem(sh=1, psv=1, nt=170, explosive=False):
    """Three-step tiles both ways (108-column strips, 8 to 20 rows), sources
    and receivers on and next to their edges (rows 8, 12, 16, 20, 24, 36,
    48, 60 start tiles of one height or another)."""
    P = inv.problems.tiny(nx=70, nz=256, nt=nt, nsrc=4, nrec=22, sh=sh, psv=psv)
    P = P.with_(src_ix=np.array([11, 16, 24, 40], np.int32),
                src_iz=np.array([107, 108, 216, 215], np.int32),
                rec_ix=np.array([4, 7, 8, 12, 15, 16, 19, 20, 23, 24, 31, 32, 35, 36, 39, 40, 47,
                                 48, 59, 60, 66, 69], np.int32),
                rec_iz=np.array([107, 108, 109, 215, 216, 217, 3, 100, 250, 60, 2, 253, 107,
                                 108, 1, 254, 216, 107, 109, 215, 120, 0], np.int32))
    if explosive:
        P = P.with_(src_type=np.array([1, 0, 1, 1], np.int32))
    return P


def _forward_all(ctx, P, batch):
    out = []
    for first in range(0, P.nsrc, batch):
        ctx.forward(first, batch, snapshots=True)
        for s in range(batch):
            out.append((ctx.traces_get(s), ctx.state_get(s), ctx.snapshots_get(s)))
    return out


@pytest.mark.parametrize("trip", TRIPS)
@pytest.mark.parametrize("pair", [None, (20, 2)], ids=["no-pair", "pair20x2"])
@pytest.mark.parametrize("case", [dict(sh=1, psv=0), dict(sh=0, psv=1), dict(sh=1, psv=1),
                                  dict(sh=0, psv=1, explosive=True)],
                         ids=["sh", "psv", "sh+psv", "psv-explosive"])
def test_triple_forward_matches_oracle(gpu, case, pair, trip):
    P = edge_problem(**case)
    print(f"triple forward {case} pair={pair} trip={trip}", flush=True)
    ctx = P.context(batch=2)
    ctx.set_math(H.INV_MATH_EXACT)
    if pair:
        ctx.set_pair(*pair)
    ctx.set_triple(trip)
    ctx.set_graphs(False)
    ctx.prof_enable(1)
    got = _forward_all(ctx, P, 2)
    n, _ = ctx.prof_read(H.INV_PROF_TRIPLE)
    ctx.close()
    assert n > 0, "three-step launches ran"
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


def _gradient(P, obs, batch, trip, fmt, graphs=True, math=H.INV_MATH_FAST, pair=(20, 2),
              misfit=0):
    ctx = P.context(batch=batch)
    ctx.set_math(math)
    ctx.set_graphs(graphs)
    ctx.set_snapshots(fmt, 0)
    if pair:
        ctx.set_pair(*pair)
    ctx.set_triple(trip)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=misfit, sigma=0)
    k = ctx.kernels_get()
    ctx.forward(0, batch, snapshots=False)
    tr = [ctx.traces_get(b) for b in range(batch)]
    st = [ctx.state_get(b) for b in range(batch)]
    ctx.close()
    return f, k, tr, st


@pytest.mark.parametrize("trip", TRIPS)
@pytest.mark.parametrize("waves", ["psv", "sh", "sh+psv"])
@pytest.mark.parametrize("batch,graphs,math,pair",
                         [(3, True, H.INV_MATH_FAST, (20, 2)), (1, False, H.INV_MATH_FAST, (8, 2)),
                          (3, False, H.INV_MATH_EXACT, None)],
                         ids=["b3-graphs-fast-pair20", "b1-direct-fast-pair8",
                              "b3-direct-exact-nopair"])
def test_triple_gradient_equals_without(gpu, waves, trip, batch, graphs, math, pair):
    sh, psv = {"psv": (0, 1), "sh": (1, 0), "sh+psv": (1, 1)}[waves]
    P = edge_problem(sh=sh, psv=psv, nt=300, explosive=bool(psv))
    P = inv.problems.shots(P, np.arange(batch))
    obs = _obs(P)
    print(f"triple gradient {waves} trip={trip} batch={batch} graphs={graphs} math={math} "
          f"pair={pair}", flush=True)
    f1, k1, t1, s1 = _gradient(P, obs, batch, trip, H.INV_SNAP_F32, graphs, math, pair)
    f0, k0, t0, s0 = _gradient(P, obs, batch, 0, H.INV_SNAP_F32, graphs, math, pair)
    assert f1 == f0
    assert np.abs(k0[1]).max() > 0 and np.abs(k0[2]).max() > 0
    for a, b in zip(k1, k0):
        assert np.array_equal(a, b), f"kernels differ, rel L2 {rel(a, b):.3e}"
    for a, b in zip(t1 + s1, t0 + s0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("fmt", [H.INV_SNAP_I16, H.INV_SNAP_F16], ids=["i16", "f16"])
def test_triple_gradient_16bit_snapshots(gpu, fmt):
    P = edge_problem(sh=0, psv=1, nt=300, explosive=True)
    P = inv.problems.shots(P, [0, 1])
    obs = _obs(P)
    print(f"triple gradient fmt={fmt}", flush=True)
    f1, k1, t1, s1 = _gradient(P, obs, 2, 12, fmt)
    f0, k0, t0, s0 = _gradient(P, obs, 2, 0, fmt)
    assert f1 == f0
    for a, b in zip(k1, k0):
        print(f"  kernel difference {rel(a, b):.2e}")
        assert rel(a, b) <= 1e-4
    for a, b in zip(t1 + s1, t0 + s0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("chunks", [[1, 40, 41, 88], [7, 7, 7, 149], [2, 5, 163]],
                         ids=["odd-first", "sevens", "short"])
def test_triple_chunked_runs_equal_whole(gpu, chunks):
    """Forward and adjoint advanced in chunks (three-step launches inside
    each chunk, the remainder in shorter launches) equal whole runs."""
    P = edge_problem(sh=0, psv=1, nt=170, explosive=True)
    P = inv.problems.shots(P, [0, 1])
    print(f"triple chunks {chunks}", flush=True)

    def run(chunked):
        ctx = P.context(batch=2)
        ctx.set_pair(20, 2)
        ctx.set_triple(12)
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
        ctx.close()
        return tr, st, k

    a, b = run(True), run(False)
    for x, y in zip(a[0] + a[1] + list(a[2]), b[0] + b[1] + list(b[2])):
        assert np.array_equal(x, y)


def test_triple_c4_grid_equals_pair(gpu):
    """The 4096x2048 grid, one shot, 60 steps with snapshots: traces,
    wavefields and snapshots with three-step launches equal the two-step
    path's bit for bit."""
    P = inv.problems.shots(inv.problems.workload("c4", nt=60, sigma=0)[0], [100])
    print(f"triple c4 {P.nx}x{P.nz} nt={P.nt}", flush=True)

    def run(trip):
        ctx = P.context(batch=1)
        ctx.set_snapshots(H.INV_SNAP_F32, 0)
        ctx.set_pair(20, 2)
        ctx.set_triple(trip)
        ctx.forward(0, 1, snapshots=True)
        out = (ctx.traces_get(0), ctx.state_get(0), ctx.snapshots_get(0))
        ctx.close()
        return out

    for x, y in zip(run(12), run(0)):
        assert np.array_equal(x, y)


def test_triple_rejects_what_it_cannot_run(gpu):
    P = edge_problem(sh=0, psv=1, nt=20)
    ctx = P.context(batch=1)
    with pytest.raises(H.InvError, match="8, 12, 16 or 20 rows"):
        ctx.set_triple(24)
    ctx.close()
    P = inv.problems.tiny(nx=40, nz=41, nt=20, nsrc=1, nrec=2, sh=0, psv=1)
    ctx = P.context(batch=1)
    with pytest.raises(H.InvError, match="nz even"):
        ctx.set_triple(12)
    ctx.close()

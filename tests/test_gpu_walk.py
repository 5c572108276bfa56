"""The walking three-step kernel (inv_set_walk): a workgroup walks down a
chunk of rows of one column strip, three time steps per launch, every step's
rows kept in LDS.  No counterpart in the reference, which launches every half
step (solver/fdm.h:534-551).  Bars:

  * exact math: traces, final wavefields and snapshots equal the FP32
    oracle's value for value (SH, P-SV, both; point forces and explosions;
    sources and receivers on and next to chunk edges; chunks of 16 to 64
    rows, a last chunk cut short by the grid's end, one chunk for the grid);
  * any math mode: gradients and a following forward equal those of the
    three-step tiles bit for bit (graph replays and direct launches, 1 and 3
    shots per launch, with and without a two-step tile), whole and chunked
    runs;
  * the 4096x2048 grid: wavefields and traces equal the two-step path's;
  * unsupported chunks and odd nz are rejected.
"""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel
from test_gpu_triple import _forward_all, _obs, edge_problem

pytestmark = pytest.mark.gpu

H = inv.invhip
WALKS = [16, 24, 40, 64, 256]


@pytest.mark.parametrize("walk", WALKS)
@pytest.mark.parametrize("case", [dict(sh=1, psv=0), dict(sh=0, psv=1), dict(sh=1, psv=1),
                                  dict(sh=0, psv=1, explosive=True)],
                         ids=["sh", "psv", "sh+psv", "psv-explosive"])
def test_walk_forward_matches_oracle(gpu, case, walk):
    P = edge_problem(**case)
    print(f"walk forward {case} walk={walk}", flush=True)
    ctx = P.context(batch=2)
    ctx.set_math(H.INV_MATH_EXACT)
    ctx.set_pair(20, 2)
    ctx.set_walk(walk)
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


def _gradient(P, obs, batch, walk, trip, graphs, math, pair):
    ctx = P.context(batch=batch)
    ctx.set_math(math)
    ctx.set_graphs(graphs)
    ctx.set_snapshots(H.INV_SNAP_F32, 0)
    if pair:
        ctx.set_pair(*pair)
    ctx.set_triple(trip)
    ctx.set_walk(walk)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=0, sigma=0)
    k = ctx.kernels_get()
    ctx.forward(0, batch, snapshots=False)
    tr = [ctx.traces_get(b) for b in range(batch)]
    st = [ctx.state_get(b) for b in range(batch)]
    ctx.close()
    return f, k, tr, st


@pytest.mark.parametrize("walk", [16, 40, 128])
@pytest.mark.parametrize("waves", ["psv", "sh", "sh+psv"])
@pytest.mark.parametrize("batch,graphs,math,pair",
                         [(3, True, H.INV_MATH_FAST, (20, 2)), (1, False, H.INV_MATH_FAST, (8, 2)),
                          (3, False, H.INV_MATH_EXACT, None)],
                         ids=["b3-graphs-fast-pair20", "b1-direct-fast-pair8",
                              "b3-direct-exact-nopair"])
def test_walk_gradient_equals_tiles(gpu, waves, walk, batch, graphs, math, pair):
    sh, psv = {"psv": (0, 1), "sh": (1, 0), "sh+psv": (1, 1)}[waves]
    P = edge_problem(sh=sh, psv=psv, nt=300, explosive=bool(psv))
    P = inv.problems.shots(P, np.arange(batch))
    obs = _obs(P)
    print(f"walk gradient {waves} walk={walk} batch={batch} graphs={graphs} math={math} "
          f"pair={pair}", flush=True)
    f1, k1, t1, s1 = _gradient(P, obs, batch, walk, 0, graphs, math, pair)
    f0, k0, t0, s0 = _gradient(P, obs, batch, 0, 12, graphs, math, pair)
    assert f1 == f0
    assert np.abs(k0[1]).max() > 0 and np.abs(k0[2]).max() > 0
    for a, b in zip(k1, k0):
        assert np.array_equal(a, b), f"kernels differ, rel L2 {rel(a, b):.3e}"
    for a, b in zip(t1 + s1, t0 + s0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("chunks", [[1, 40, 41, 88], [7, 7, 7, 149]], ids=["odd-first", "sevens"])
def test_walk_chunked_runs_equal_whole(gpu, chunks):
    P = edge_problem(sh=0, psv=1, nt=170, explosive=True)
    P = inv.problems.shots(P, [0, 1])
    print(f"walk chunks {chunks}", flush=True)

    def run(chunked):
        ctx = P.context(batch=2)
        ctx.set_pair(20, 2)
        ctx.set_walk(24)
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


def test_walk_c4_grid_equals_pair(gpu):
    """The 4096x2048 grid, one shot, 60 steps with snapshots, chunks of 256
    rows: traces, wavefields and snapshots equal the two-step path's."""
    P = inv.problems.shots(inv.problems.workload("c4", nt=60, sigma=0)[0], [100])
    print(f"walk c4 {P.nx}x{P.nz} nt={P.nt}", flush=True)

    def run(walk):
        ctx = P.context(batch=1)
        ctx.set_snapshots(H.INV_SNAP_F32, 0)
        ctx.set_pair(20, 2)
        ctx.set_walk(walk)
        ctx.forward(0, 1, snapshots=True)
        out = (ctx.traces_get(0), ctx.state_get(0), ctx.snapshots_get(0))
        ctx.close()
        return out

    for x, y in zip(run(256), run(0)):
        assert np.array_equal(x, y)


def test_walk_rejects_what_it_cannot_run(gpu):
    P = edge_problem(sh=0, psv=1, nt=20)
    ctx = P.context(batch=1)
    for bad in (8, 20, 4104, -8):
        with pytest.raises(H.InvError, match="multiple of 8"):
            ctx.set_walk(bad)
    ctx.close()
    P = inv.problems.tiny(nx=40, nz=41, nt=20, nsrc=1, nrec=2, sh=0, psv=1)
    ctx = P.context(batch=1)
    with pytest.raises(H.InvError, match="nz even"):
        ctx.set_walk(64)
    ctx.close()

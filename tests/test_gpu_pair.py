"""The two-step kernel (inv_set_pair): two time steps per launch with the
tile's stress and velocity kept in LDS between them.  No counterpart in the
reference, which launches every half step (solver/fdm.h:534-551).  Bars:

  * exact math: traces, final wavefields and snapshots equal the FP32
    oracle's value for value (SH, P-SV, both; point forces and explosions;
    sources and receivers on the pair tiles' row and column edges);
  * any math mode: runs with the two-step kernel give bitwise the results of
    the one-step kernel for forward runs and FP32-snapshot gradients, with
    checkpoint recomputation, graph replays and direct launches, whole and
    chunked runs; with 16-bit snapshots (scaled per row segment, whose width
    is the writing tile's: 116 columns against 120) the gradients agree to
    the format's tolerance;
  * unsupported tiles are rejected with an error.
"""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel

pytestmark = pytest.mark.gpu

H = inv.invhip
PAIRS = [(4, 2), (8, 2), (12, 2), (16, 2), (20, 2), (4, 4), (8, 4)]


def _pid(t):
    return f"pair{t[0]}x{t[1]}"


def edge_problem(sh=1, psv=1, nt=170, explosive=False):
    """Several two-step tiles both ways (strips of 116 or 240 columns, 4-12
    rows), sources and receivers on and next to their edges."""
    P = inv.problems.tiny(nx=70, nz=256, nt=nt, nsrc=4, nrec=12, sh=sh, psv=psv)
    P = P.with_(src_ix=np.array([11, 12, 24, 36], np.int32),
                src_iz=np.array([115, 116, 240, 232], np.int32),
                rec_ix=np.array([4, 8, 12, 13, 23, 24, 35, 36, 47, 48, 60, 66], np.int32),
                rec_iz=np.array([117, 116, 115, 231, 232, 239, 240, 241, 3, 120, 250, 60],
                                np.int32))
    if explosive:
        P = P.with_(src_type=np.array([1, 0, 1, 1], np.int32))
    return P


def _pair(ctx, tile):
    try:
        ctx.set_pair(*tile)
    except H.InvError:
        ctx.close()
        raise


def _forward_all(ctx, P, batch):
    out = []
    for first in range(0, P.nsrc, batch):
        ctx.forward(first, batch, snapshots=True)
        for s in range(batch):
            out.append((ctx.traces_get(s), ctx.state_get(s), ctx.snapshots_get(s)))
    return out


@pytest.mark.parametrize("tile", PAIRS, ids=_pid)
@pytest.mark.parametrize("case", [dict(sh=1, psv=0), dict(sh=0, psv=1), dict(sh=1, psv=1),
                                  dict(sh=0, psv=1, explosive=True)],
                         ids=["sh", "psv", "sh+psv", "psv-explosive"])
def test_pair_forward_matches_oracle(gpu, case, tile):
    P = edge_problem(**case)
    print(f"pair forward {case} {_pid(tile)}", flush=True)
    ctx = P.context(batch=2)
    ctx.set_math(H.INV_MATH_EXACT)
    try:
        _pair(ctx, tile)
        got = _forward_all(ctx, P, 2)
    except H.InvError as e:
        pytest.skip(f"tile not supported here: {e}")
    ctx.close()
    O = oracle_problem(P)
    for s in range(P.nsrc):
        otr, osn, ost = O.forward(s, snapshots=True)
        tr, st, sn = got[s]
        assert np.array_equal(tr, oracle_traces(otr, P)), f"traces of shot {s}"
        assert np.array_equal(st, ost), f"wavefields of shot {s}"
        assert np.array_equal(sn, osn), f"snapshots of shot {s}"


def _gradient(P, obs, batch, tile, fmt, ckpt, graphs=True, misfit=0, math=H.INV_MATH_FAST):
    ctx = P.context(batch=batch)
    ctx.set_math(math)
    ctx.set_graphs(graphs)
    ctx.set_snapshots(fmt, ckpt)
    if tile:
        ctx.set_pair(*tile)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=misfit, sigma=0)
    k = ctx.kernels_get()
    ctx.forward(0, batch, snapshots=False)
    tr = [ctx.traces_get(b) for b in range(batch)]
    st = [ctx.state_get(b) for b in range(batch)]
    ctx.close()
    return f, k, tr, st


def _obs(P):
    O = oracle_problem(P, model="true")
    return np.stack([oracle_traces(O.forward(s)[0], P) for s in range(P.nsrc)]).astype(np.float32)


SEG_TOL = {H.INV_SNAP_F32: 0.0, H.INV_SNAP_F16: 1e-6, H.INV_SNAP_BF16: 1e-6,
           H.INV_SNAP_I16: 1e-4}


@pytest.mark.parametrize("tile", [(8, 2), (4, 4), (12, 2), (16, 2), (20, 2)], ids=_pid)
@pytest.mark.parametrize("fmt,ckpt,graphs", [(H.INV_SNAP_F32, 0, True), (H.INV_SNAP_F16, 100, True),
                                             (H.INV_SNAP_BF16, 60, False),
                                             (H.INV_SNAP_F16, 0, False),
                                             (H.INV_SNAP_I16, 100, True)],
                         ids=["f32", "f16-ckpt", "bf16-ckpt-direct", "f16-direct", "i16-ckpt"])
def test_pair_gradient_equals_one_step(gpu, tile, fmt, ckpt, graphs):
    """Gradient (forward with snapshots, misfit, adjoint, recomputed segments)
    and a following forward: the one-step kernel's, bit for bit except the
    16-bit kernels (SEG_TOL)."""
    P = edge_problem(sh=1, psv=1, nt=300, explosive=True)
    obs = _obs(P)
    print(f"pair gradient {_pid(tile)} fmt={fmt} ckpt={ckpt} graphs={graphs}", flush=True)
    try:
        f1, k1, t1, s1 = _gradient(P, obs, 2, tile, fmt, ckpt, graphs)
    except H.InvError as e:
        pytest.skip(f"tile not supported here: {e}")
    f0, k0, t0, s0 = _gradient(P, obs, 2, None, fmt, ckpt, graphs)
    assert f1 == f0
    for a, b in zip(k1, k0):
        assert np.abs(b).max() > 0
        if fmt == H.INV_SNAP_F32:
            assert np.array_equal(a, b), f"kernels differ, rel L2 {rel(a, b):.3e}"
        else:
            print(f"  kernel difference {rel(a, b):.2e}")
            assert rel(a, b) <= SEG_TOL[fmt]
    for a, b in zip(t1 + s1, t0 + s0):
        assert np.array_equal(a, b)


def test_pair_gradient_matches_oracle(gpu):
    """One shot per launch, exact math: the gradient equals the FP32
    oracle's value for value."""
    P = edge_problem(sh=1, psv=1, nt=200, explosive=True)
    obs = _obs(P)
    f, k, _, _ = _gradient(P, obs, 1, (8, 2), H.INV_SNAP_F32, 0, math=H.INV_MATH_EXACT)
    O = oracle_problem(P)
    f32, *k32 = O.gradient(obs, 0, 0, dtype=np.float32)
    assert abs(f - f32) / abs(f32) < 1e-6
    for a, b, name in zip(k, k32, ("k_lambda", "k_mu", "k_rho")):
        assert np.array_equal(a, b), f"{name} rel L2 {rel(a, b):.3e}"


@pytest.mark.parametrize("odd", [False, True], ids=["even-nt", "odd-chunks"])
def test_pair_chunked_runs_equal_whole(gpu, odd):
    """Runs advanced in chunks (pairs break at chunk ends) equal whole runs,
    forward and adjoint, with FP16 snapshots and checkpoints."""
    P = edge_problem(sh=0, psv=1, nt=300)
    ctx = P.context(batch=2)
    ctx.set_snapshots(H.INV_SNAP_F16, 60)
    ctx.set_pair(8, 2)
    rng = np.random.default_rng(7)
    ads = rng.standard_normal((2, P.ncomp, P.nrec, P.nt)).astype(np.float32)
    ctx.forward(0, 2, snapshots=True)
    whole_tr = [ctx.traces_get(b) for b in range(2)]
    for b in range(2):
        ctx.adstf_set(b, ads[b])
    ctx.kernels_zero()
    ctx.adjoint(0, 2)
    whole_k = ctx.kernels_get()
    step = 37 if odd else 50
    ctx.forward_begin(0, 2, snapshots=True)
    pos = 0
    while pos < P.nt:
        pos = ctx.advance(step)
    chunk_tr = [ctx.traces_get(b) for b in range(2)]
    for b in range(2):
        ctx.adstf_set(b, ads[b])
    ctx.kernels_zero()
    ctx.adjoint_begin(0, 2)
    pos = 0
    while pos < P.nt:
        pos = ctx.advance(step)
    chunk_k = ctx.kernels_get()
    ctx.close()
    for a, b in zip(whole_tr + list(whole_k), chunk_tr + list(chunk_k)):
        assert np.array_equal(a, b)


def test_pair_profile_counts_launches(gpu):
    """Sampled two-step launches are counted under INV_PROF_PAIR."""
    P = edge_problem(sh=0, psv=1, nt=200)
    ctx = P.context(batch=1)
    ctx.set_graphs(False)
    ctx.set_pair(8, 2)
    ctx.prof_enable(1)
    ctx.forward(0, 1, snapshots=False)
    n, ms = ctx.prof_read(H.INV_PROF_PAIR)
    ctx.close()
    assert n == P.nt // 2 and ms > 0


def test_pair_rejects_bad_tiles(gpu):
    P = inv.problems.tiny(nx=40, nz=62, nt=20, sh=0, psv=1)  # nz not divisible by 4
    ctx = P.context(batch=1)
    for t in ((24, 2), (6, 2), (12, 4), (8, 1), (8, 3)):
        with pytest.raises(H.InvError):
            ctx.set_pair(*t)
    with pytest.raises(H.InvError):
        ctx.set_pair(8, 4)
    ctx.set_pair(0, 2)  # off
    ctx.close()


@pytest.mark.parametrize("math", [0, 1], ids=["exact", "fast"])
@pytest.mark.parametrize("explosive", [False, True], ids=["force", "explosive"])
def test_every_kernel_records_the_same(gpu, math, explosive):
    """Traces, wavefields and snapshots of the one-column, wide and two-step
    kernels agree bit for bit in both math modes.  Every tile here holds a
    receiver or source as the first entry of its point list, which the
    kernels keep per lane and read with lane reads: a scan from code where
    that lane is inactive lost such entries."""
    P = edge_problem(sh=1, psv=1, nt=300, explosive=explosive)
    out = {}
    for name, tile, pair in (("col8", (8, 1), None), ("col16", (16, 1), None),
                             ("wide4x2", (4, 2), None), ("wide8x4", (8, 4), None),
                             ("pair8x2", (8, 1), (8, 2)), ("pair16x2", (8, 1), (16, 2)),
                             ("pair20x2", (8, 1), (20, 2)),
                             ("pair4x4", (8, 1), (4, 4))):
        ctx = P.context(batch=2)
        ctx.set_math(math)
        ctx.set_tile(*tile)
        if pair:
            ctx.set_pair(*pair)
        ctx.forward(0, 2, snapshots=True)
        out[name] = [ctx.traces_get(b) for b in range(2)] + \
                    [ctx.state_get(b) for b in range(2)] + [ctx.snapshots_get(b) for b in range(2)]
        ctx.close()
    ref = out.pop("col8")
    for name, got in out.items():
        for k, (a, b) in enumerate(zip(got, ref)):
            assert np.array_equal(a, b), f"{name} differs from col8 in array {k}"


def _run(P, tile, pair, fmt, ckpt, obs=None, snapshots=True):
    ctx = P.context(batch=2)
    ctx.set_snapshots(fmt, ckpt)
    ctx.set_tile(*tile)
    if pair:
        ctx.set_pair(*pair)
    if obs is None:
        ctx.forward(0, 2, snapshots=snapshots)
        out = [ctx.traces_get(b) for b in range(2)] + [ctx.state_get(b) for b in range(2)]
    else:
        for s in range(P.nsrc):
            ctx.obs_set(s, obs[s])
        out = [ctx.gradient(misfit_type=0, sigma=0)] + list(ctx.kernels_get())
    ctx.close()
    return out


@pytest.mark.parametrize("tile,pair", [((8, 2), None), ((8, 1), (16, 2)), ((8, 1), (20, 2))],
                         ids=["wide8x2", "pair16x2", "pair20x2"])
def test_kernel_variants_agree(gpu, tile, pair):
    """The step kernels are compiled with and without snapshot roles and
    receiver recording, picked per launch.  A forward that stores snapshots
    gives the same traces and wavefields as one that does not (FP32 and
    FP16), and a gradient whose adjoint recomputes checkpoint segments
    (launches without recording) equals the one that keeps every snapshot."""
    P = edge_problem(sh=1, psv=1, nt=300, explosive=True)
    print(f"variants {tile} {pair}", flush=True)
    for fmt in (H.INV_SNAP_F32, H.INV_SNAP_F16):
        a = _run(P, tile, pair, fmt, 0, snapshots=True)
        b = _run(P, tile, pair, fmt, 0, snapshots=False)
        for k, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x, y), f"fmt {fmt}: array {k} depends on snapshot storage"
    obs = _obs(P)
    whole = _run(P, tile, pair, H.INV_SNAP_F16, 0, obs=obs)
    ckpt = _run(P, tile, pair, H.INV_SNAP_F16, 60, obs=obs)
    assert whole[0] == ckpt[0]
    for x, y in zip(whole[1:], ckpt[1:]):
        assert np.abs(x).max() > 0
        assert np.array_equal(x, y)

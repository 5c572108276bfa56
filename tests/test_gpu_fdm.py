"""GPU parity of the propagation path (through the C ABI) against the oracle.

Bars: in exact math mode the FP32 results equal the FP32 oracle's (same
expression order, IEEE divisions, no fused multiply-add on either side); in
fast math mode (coefficient multiplies) every FP32 wavefield is within 1e-5
relative L2 of the FP32 and of the FP64 oracle.  Kernels within 1e-4.
"""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel

pytestmark = pytest.mark.gpu

CASES = [dict(sh=1, psv=0), dict(sh=0, psv=1), dict(sh=1, psv=1)]
# (variant, tile, math): fused column kernel (0) with tile_x rows, or with a
# (rows, columns per lane) tile of the wide kernel; velocity/stress kernel
# pair (1); fused linear-index kernel (2); in exact (0) and fast (1) math
VARIANTS = [(0, 8, 0), (0, 16, 0), (0, 32, 0), (1, 0, 0), (2, 8, 0),
            (0, (4, 2), 0), (0, (16, 2), 0), (0, (4, 4), 0), (0, (8, 4), 0),
            (0, 8, 1), (0, 16, 1), (1, 0, 1), (2, 16, 1), (0, (8, 2), 1), (0, (4, 4), 1),
            (0, (8, 4), 1)]


def _vid(v):
    if isinstance(v[1], tuple):
        kind = f"col{v[1][0]}x{v[1][1]}"
    else:
        kind = {0: f"col{v[1]}", 1: "split", 2: f"lin{v[1]}"}[v[0]]
    return kind + ("-fast" if v[2] else "-exact")


def setup(ctx, v):
    if isinstance(v[1], tuple):
        if ctx.desc.nz % v[1][1]:
            ctx.close()
            pytest.skip(f"nz={ctx.desc.nz} is not divisible by {v[1][1]} columns per lane")
        ctx.set_tile(*v[1])
    else:
        ctx.set_variant(v[0], v[1])
    ctx.set_math(v[2])


def same(a, ref, fast, tol=1e-5):
    """Exact mode: equal values.  Fast mode: within tol relative L2."""
    if fast:
        e = rel(a, ref)
        assert e < tol, f"rel L2 {e:.3e} >= {tol}"
    else:
        assert np.array_equal(a, ref), f"not equal, rel L2 {rel(a, ref):.3e}"


def boundary_problem(**kw):
    """A grid with several tiles in both directions and sources/receivers on
    tile edges: rows 15/16/31/32 (tiles of 4, 8 and 16 rows); columns 57/58
    and 115/116 (58-column tiles, one column per lane), 119/120 and 239/240
    (120 columns, two per lane), 247/248/249 (248 columns, four per lane),
    63/64/127/128 (64-column tiles).  nz = 256 is divisible by 4, so every
    tile runs and the four-column tiles have a z-tile edge."""
    P = inv.problems.tiny(nx=70, nz=256, nt=160, nsrc=4, nrec=14, **kw)
    src_ix = np.array([15, 16, 31, 32], np.int32)
    src_iz = np.array([64, 116, 240, 248], np.int32)
    rec_ix = np.array([8, 15, 16, 17, 31, 32, 33, 47, 48, 60, 16, 31, 40, 63], np.int32)
    rec_iz = np.array([57, 58, 63, 64, 119, 120, 127, 247, 248, 249, 115, 239, 3, 252],
                      np.int32)
    return P.with_(src_ix=src_ix, src_iz=src_iz, rec_ix=rec_ix, rec_iz=rec_iz)


def _ids(c):
    return "sh" if c["sh"] and not c["psv"] else ("psv" if not c["sh"] else "sh+psv")


@pytest.mark.parametrize("variant", VARIANTS, ids=_vid)
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_forward_plain_matches_oracle(gpu, case, variant):
    P = inv.problems.tiny(**case)
    print(f"forward plain {_ids(case)} {_vid(variant)} {P.nx}x{P.nz} nt={P.nt}", flush=True)
    ctx = P.context(batch=1)
    setup(ctx, variant)
    fast = variant[2]
    O = oracle_problem(P)
    for isrc in range(P.nsrc):
        ctx.forward(isrc, 1, snapshots=True)
        tr = ctx.traces_get(0)
        st = ctx.state_get(0)
        sn = ctx.snapshots_get(0)
        otr, osn, ost = O.forward(isrc, snapshots=True)
        ref = oracle_traces(otr, P)
        same(tr, ref, fast)
        same(st, ost, fast)
        same(sn, osn, fast)
        o64, _, s64 = O.forward(isrc, dtype=np.float64)
        assert rel(tr, oracle_traces(o64, P)) < 1e-5
        assert rel(st, s64) < 1e-5
    ctx.close()


def test_forward_all_sources_and_velocity_traces(gpu):
    """runForward(-1) fires every source; adj = 0 with trace_type 0 records v."""
    P = inv.problems.tiny(sh=1, psv=1)
    ctx = P.context(batch=1, adj=0)
    ctx.set_math(0)
    O = oracle_problem(P, adj=0)
    ctx.forward(-1, 1, snapshots=False)
    tr = ctx.traces_get(0)
    otr, _, _ = O.forward(-1)
    assert np.array_equal(tr, oracle_traces(otr, P))
    ctx.close()


def test_displacement_traces(gpu):
    P = inv.problems.tiny(sh=0, psv=1).with_(trace_type=1)
    ctx = P.context(batch=1, adj=0)
    ctx.set_math(0)
    O = oracle_problem(P, adj=0)
    ctx.forward(1, 1, snapshots=False)
    otr, _, _ = O.forward(1)
    assert np.array_equal(ctx.traces_get(0), oracle_traces(otr, P))
    ctx.close()


@pytest.mark.parametrize("variant", VARIANTS, ids=_vid)
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_tile_edges_match_oracle(gpu, case, variant):
    """Sources and receivers on tile edges and in tile halos, several tiles."""
    P = boundary_problem(**case)
    print(f"tile edges {_ids(case)} {_vid(variant)}", flush=True)
    ctx = P.context(batch=2)
    setup(ctx, variant)
    fast = variant[2]
    O = oracle_problem(P)
    ctx.forward(1, 2, snapshots=True)
    for s in range(2):
        otr, osn, ost = O.forward(1 + s, snapshots=True)
        same(ctx.traces_get(s), oracle_traces(otr, P), fast)
        same(ctx.state_get(s), ost, fast)
        same(ctx.snapshots_get(s), osn, fast)
    ctx.close()


@pytest.mark.parametrize("variant", VARIANTS, ids=_vid)
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_batched_forward_equals_plain(gpu, case, variant):
    """Each slot of a batched launch gives the same bits as the one-shot path."""
    P = inv.problems.tiny(nsrc=4, **case)
    plain = P.context(batch=1)
    batch = P.context(batch=4)
    setup(plain, variant)
    setup(batch, variant)
    batch.forward(0, 4, snapshots=True)
    for s in range(4):
        plain.forward(s, 1, snapshots=True)
        assert np.array_equal(batch.traces_get(s), plain.traces_get(0))
        assert np.array_equal(batch.state_get(s), plain.state_get(0))
        assert np.array_equal(batch.snapshots_get(s), plain.snapshots_get(0))
    plain.close()
    batch.close()


@pytest.mark.parametrize("variant", VARIANTS, ids=_vid)
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_adjoint_plain_matches_oracle(gpu, case, variant):
    P = boundary_problem(**case)
    ctx = P.context(batch=1)
    setup(ctx, variant)
    fast = variant[2]
    O = oracle_problem(P)
    ctx.kernels_zero()
    ks = [np.zeros(P.N, np.float32) for _ in range(3)]
    for isrc in range(P.nsrc):
        ctx.forward(isrc, 1, snapshots=True)
        tr = ctx.traces_get(0)
        adstf = np.ascontiguousarray(tr[:, :, ::-1] * np.float32(0.5))
        ctx.adstf_set(0, adstf)
        ctx.adjoint(isrc, 1)
        _, osn, _ = O.forward(isrc, snapshots=True)
        comps = (["y"] if P.sh else []) + (["x", "z"] if P.psv else [])
        if fast:
            # drive the oracle with the GPU's own snapshots: this compares the
            # adjoint alone, not the forward drift before it
            osn = ctx.snapshots_get(0)
        O.adjoint({c: adstf[i] for i, c in enumerate(comps)}, osn, kernels=ks)
        same(ctx.state_get(0), O.adjoint({c: adstf[i] for i, c in enumerate(comps)}, osn)[3],
             fast)
    kl, km, kr = ctx.kernels_get()
    for a, b, name in ((kl, ks[0], "k_lambda"), (km, ks[1], "k_mu"), (kr, ks[2], "k_rho")):
        if not P.psv and name == "k_lambda":
            assert not np.any(a)
            continue
        same(a, b, fast, tol=1e-4)
    ctx.close()


def test_autotune_selects_a_candidate_and_keeps_results(gpu):
    """inv_autotune picks one of its candidates, the results after it equal
    the column kernel's bit for bit, and a second context of the same shape
    gets the cached choice."""
    P = inv.problems.tiny(sh=1, psv=1, nsrc=3, nx=70, nz=120, nt=120)
    ref = P.context(batch=3)
    ref.forward(0, 3, snapshots=False)
    want = [ref.traces_get(b) for b in range(3)]
    ref.close()
    ctx = P.context(batch=3)
    choice = ctx.autotune(0, 3, steps=10)
    print(f"autotune choice {choice}", flush=True)
    assert choice[0] == 0
    assert (choice[1], choice[2]) in [(8, 1), (16, 1), (32, 1), (4, 2), (8, 2), (16, 2), (4, 4),
                                      (8, 4)]
    assert P.nz % choice[2] == 0
    ctx.forward(0, 3, snapshots=False)
    for b in range(3):
        assert np.array_equal(ctx.traces_get(b), want[b])
    ctx.close()
    again = P.context(batch=3)
    assert again.autotune(0, 3, steps=10) == choice
    again.close()

"""GPU parity of explosive sources (source type 1: the wavelet's y trace
added to sxx and szz after the stress update, P-SV only; no counterpart in the
reference, whose sources are point forces, solver/fdm.h:163-179).  The oracle
pins the injection by one step from rest and by the x <-> z symmetry of an
explosion at the centre of a square grid; here the HIP kernels are held to
the oracle's bar for point forces: equal values in exact math, 1e-5 relative
L2 in fast math, with point forces and explosions mixed in one launch and
sources on tile edges and halos."""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel

pytestmark = pytest.mark.gpu

H = inv.invhip
TILES = [(8, 0), (16, 0), ((4, 2), 0), ((16, 2), 0), ((4, 4), 0), ((8, 4), 0),
         (8, 1), ((8, 2), 1), ((4, 4), 1)]


def _tid(t):
    kind = f"col{t[0][0]}x{t[0][1]}" if isinstance(t[0], tuple) else f"col{t[0]}"
    return kind + ("-fast" if t[1] else "-exact")


def mixed_problem(sh=1, nt=160):
    """Several tiles both ways, sources on tile edges (rows 15/16/31, columns
    63/64/120/128), types explosive, force, explosive, explosive."""
    P = inv.problems.tiny(nx=70, nz=152, nt=nt, nsrc=4, nrec=10, sh=sh, psv=1)
    return P.with_(src_ix=np.array([15, 16, 31, 32], np.int32),
                   src_iz=np.array([63, 64, 120, 128], np.int32),
                   rec_ix=np.array([8, 15, 16, 17, 31, 32, 33, 47, 48, 60], np.int32),
                   rec_iz=np.array([63, 64, 65, 127, 128, 3, 64, 63, 120, 146], np.int32),
                   src_type=np.array([1, 0, 1, 1], np.int32))


def _setup(ctx, t):
    if isinstance(t[0], tuple):
        ctx.set_tile(*t[0])
    else:
        ctx.set_variant(H.INV_VARIANT_FUSED, t[0])
    ctx.set_math(t[1])


def _same(a, ref, fast, tol=1e-5):
    if fast:
        e = rel(a, ref)
        assert e < tol, f"rel L2 {e:.3e} >= {tol}"
    else:
        assert np.array_equal(a, ref), f"not equal, rel L2 {rel(a, ref):.3e}"


@pytest.mark.parametrize("tile", TILES, ids=_tid)
@pytest.mark.parametrize("sh", [0, 1], ids=["psv", "sh+psv"])
def test_explosive_forward_matches_oracle(gpu, sh, tile):
    P = mixed_problem(sh=sh)
    print(f"explosive forward sh={sh} {_tid(tile)}", flush=True)
    ctx = P.context(batch=2)
    _setup(ctx, tile)
    O = oracle_problem(P)
    for first in (0, 2):
        ctx.forward(first, 2, snapshots=True)
        for s in range(2):
            otr, osn, ost = O.forward(first + s, snapshots=True)
            _same(ctx.traces_get(s), oracle_traces(otr, P), tile[1])
            _same(ctx.state_get(s), ost, tile[1])
            _same(ctx.snapshots_get(s), osn, tile[1])
            o64, _, s64 = O.forward(first + s, dtype=np.float64)
            assert rel(ctx.traces_get(s), oracle_traces(o64, P)) < 1e-5
    ctx.close()


@pytest.mark.parametrize("tile", [(8, 0), ((8, 2), 0), ((4, 4), 1)], ids=_tid)
def test_explosive_gradient_matches_oracle(gpu, tile):
    P = mixed_problem(sh=1)
    O_true = oracle_problem(P, model="true")
    obs = np.stack([oracle_traces(O_true.forward(s)[0], P)
                    for s in range(P.nsrc)]).astype(np.float32)
    print(f"explosive gradient {_tid(tile)}", flush=True)
    ctx = P.context(batch=1)
    _setup(ctx, tile)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=0, sigma=0)
    k = ctx.kernels_get()
    ctx.close()
    O = oracle_problem(P)
    f32, *k32 = O.gradient(obs, 0, 0, dtype=np.float32)
    f64, *k64 = O.gradient(obs.astype(np.float64), 0, 0, dtype=np.float64)
    assert abs(f - f64) / abs(f64) < 1e-5
    for a, b, c, name in zip(k, k32, k64, ("k_lambda", "k_mu", "k_rho")):
        print(f"  {name}: vs f32 oracle {rel(a, b):.2e}, vs f64 oracle {rel(a, c):.2e}")
        if tile[1] == 0:
            assert np.array_equal(a, b), f"{name} not bit-exact ({rel(a, b):.3e})"
        assert rel(a, b) < 1e-4 and rel(a, c) < 1e-4


@pytest.mark.parametrize("fmt", [H.INV_SNAP_F32, H.INV_SNAP_F16])
def test_explosive_checkpointed_gradient_is_bitwise(gpu, fmt):
    """The adjoint's segment recomputation injects the explosions again: a
    checkpointed gradient equals the one that kept every snapshot."""
    P = mixed_problem(sh=0, nt=300)
    rng = np.random.default_rng(3)
    adstf = rng.standard_normal((2, P.ncomp, P.nrec, P.nt)).astype(np.float32)
    out = []
    for ckpt in (0, 60):
        ctx = P.context(batch=2)
        ctx.set_snapshots(fmt, ckpt)
        ctx.forward(0, 2, snapshots=True)
        for b in range(2):
            ctx.adstf_set(b, adstf[b])
        ctx.kernels_zero()
        ctx.adjoint(0, 2)
        out.append(ctx.kernels_get())
        ctx.close()
    for a, b in zip(*out):
        assert np.abs(a).max() > 0 and np.array_equal(a, b)


def test_explosive_rejected_outside_the_fused_kernels(gpu):
    P = mixed_problem(sh=0)
    ctx = P.context(batch=1)
    for v in (H.INV_VARIANT_SPLIT, H.INV_VARIANT_FUSED_LINEAR):
        ctx.set_variant(v)
        with pytest.raises(H.InvError):
            ctx.forward(0, 1, snapshots=False)
    ctx.set_variant(H.INV_VARIANT_FUSED)
    ctx.forward(0, 1, snapshots=False)
    with pytest.raises(H.InvError):
        ctx.sources_type(np.array([2, 0, 0, 0], np.int32))
    ctx.close()


@pytest.fixture(scope="module")
def c0():
    """The checkerboard workload: 200x200 at 10 m, 4 explosive Ricker sources
    at 20 cells depth, 40 receivers, nt 2000, sigma 5; observed data from the
    FP32 oracle on the true model."""
    P = inv.problems.checker(explosive=True)
    O = oracle_problem(P, model="true")
    obs = np.stack([oracle_traces(O.forward(s)[0], P) for s in range(P.nsrc)])
    return P, obs.astype(np.float32)


@pytest.mark.parametrize("batch,math", [(1, 0), (4, 1)], ids=["plain-exact", "batched-fast"])
def test_c0_checkerboard_gradient_matches_oracle(gpu, c0, batch, math):
    P, obs = c0
    print(f"c0 gradient batch={batch} math={math}", flush=True)
    ctx = P.context(batch=batch)
    ctx.set_math(math)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=0, sigma=P.sigma)
    k = ctx.kernels_get()
    ctx.close()
    O = oracle_problem(P)
    f32, *k32 = O.gradient(obs, 0, P.sigma, dtype=np.float32)
    f64, *k64 = O.gradient(obs.astype(np.float64), 0, P.sigma, dtype=np.float64)
    assert abs(f - f64) / abs(f64) < 1e-5
    for a, b, c, name in zip(k, k32, k64, ("k_lambda", "k_mu", "k_rho")):
        print(f"  {name}: vs f32 oracle {rel(a, b):.2e}, vs f64 oracle {rel(a, c):.2e}")
        assert np.abs(b).max() > 0
        if batch == 1 and math == 0:
            assert np.array_equal(a, b), f"{name} not bit-exact ({rel(a, b):.3e})"
        assert rel(a, b) < 1e-4 and rel(a, c) < 1e-4

"""Reduced-precision forward snapshots and checkpoint recomputation
(inv_set_snapshots).  No counterpart in the reference, which keeps FP32
snapshots of every sfe-th step (solver/fdm.h:521-560); the checks are:

  * 16-bit snapshots decode to the FP32 snapshots within the rounding bound
    of the format (FP16: 2^-11 |x|, BF16: 2^-8 |x|, plus FP16's subnormal floor
    2^-38 of the field's max-abs; I16: 2^-15 of the max-abs of the value's
    block of 4 cells, bounded by that of the 7 cells around it, or 2^-30 of
    its row's max-abs where the block offset is clamped), slot by slot and
    field by field;
  * gradients from 16-bit snapshots stay within a stated error of the FP64
    oracle (reported per format);
  * checkpointed runs (the adjoint recomputes each forward segment from a
    saved FP32 state) give bit for bit the gradient of the run that kept
    every snapshot, for every format, batch size and sfe, captured in a graph,
    launched directly and advanced in chunks;
  * unsupported combinations are rejected with an error.
"""
import numpy as np
import pytest

from conftest import inv, oracle_problem, oracle_traces, rel

pytestmark = pytest.mark.gpu

H = inv.invhip
FMTS = {"f32": H.INV_SNAP_F32, "f16": H.INV_SNAP_F16, "bf16": H.INV_SNAP_BF16,
        "i16": H.INV_SNAP_I16}
# relative L2 between 16-bit gradients of kernels whose tiles (and so row
# segments) differ: FP16 / BF16 rounding does not depend on the power-of-two
# scale except in FP16's subnormal range (below 2^-28 of a segment's max);
# I16 quantises relative to the segment max
SEG_TOL = {"f32": 0.0, "f16": 1e-6, "bf16": 1e-6, "i16": 1e-4}


def _obs(P):
    O = oracle_problem(P, model="true")
    return np.stack([oracle_traces(O.forward(s)[0], P) for s in range(P.nsrc)]).astype(np.float32)


def _gradient(P, obs, batch, fmt=0, ckpt=0, graphs=True, misfit=0, sigma=0):
    ctx = P.context(batch=batch)
    ctx.set_graphs(graphs)
    ctx.set_snapshots(fmt, ckpt)
    for s in range(P.nsrc):
        ctx.obs_set(s, obs[s])
    f = ctx.gradient(misfit_type=misfit, sigma=sigma)
    k = ctx.kernels_get()
    ctx.close()
    return f, k


@pytest.mark.parametrize("fmt", ["f16", "bf16", "i16"])
@pytest.mark.parametrize("sfe", [1, 10])
def test_snapshot_decode_bound(gpu, fmt, sfe):
    P = inv.problems.tiny(sh=1, psv=1, nsrc=2, nt=200, sfe=sfe)
    print(f"snapshots {fmt} sfe={sfe}", flush=True)
    ctx = P.context(batch=2)
    ctx.forward(0, 2, snapshots=True)
    ref = [ctx.snapshots_get(b) for b in range(2)]
    ctx.set_snapshots(FMTS[fmt], 0)
    ctx.forward(0, 2, snapshots=True)
    got = [ctx.snapshots_get(b) for b in range(2)]
    ctx.close()
    for b in range(2):
        x, y = ref[b].astype(np.float64), got[b].astype(np.float64)
        m = np.abs(x).max(axis=2, keepdims=True)
        if fmt == "i16":
            # [slots][fields][nx][nz]: a value's step is set by its block of
            # 4 cells (inside the 7 cells around it) unless the block's offset
            # is clamped at 15 below its row segment (inside its row)
            a4 = np.abs(x).reshape(x.shape[0], x.shape[1], P.nx, P.nz)
            pad = np.pad(a4, ((0, 0), (0, 0), (0, 0), (3, 3)))
            win = np.max(np.stack([pad[..., s:s + P.nz] for s in range(7)]), axis=0)
            rows = a4.max(axis=3, keepdims=True)
            # plus half a step of the smallest scale (exponents are clamped
            # to 120: blocks whose max-abs is below 2^-105)
            bound = ((win + rows * 2.0 ** -15) * 2.0 ** -15 * (1 + 1e-4)).reshape(x.shape) \
                + 2.0 ** -121
        else:
            eps = 2.0 ** -11 if fmt == "f16" else 2.0 ** -8
            bound = eps * np.abs(x) * (1 + 1e-6) + m * 2.0 ** -38
        worst = np.max(np.abs(y - x) - bound)
        assert worst <= 0, f"slot {b}: decode error exceeds the bound by {worst:.3e}"
        assert np.count_nonzero(m) > m.size // 2  # the snapshots are not trivially zero


@pytest.mark.parametrize("fmt,tol", [("f16", 2e-3), ("bf16", 2e-2), ("i16", 1e-4)])
def test_gradient_error_16bit_vs_fp64(gpu, fmt, tol):
    P = inv.problems.tiny(sh=1, psv=1, nsrc=3)
    obs = _obs(P)
    f, k = _gradient(P, obs, batch=3, fmt=FMTS[fmt])
    f32, k32 = _gradient(P, obs, batch=3, fmt=H.INV_SNAP_F32)
    O = oracle_problem(P)
    f64, *k64 = O.gradient(obs.astype(np.float64), 0, 0, dtype=np.float64)
    assert f == f32  # the misfit does not depend on the snapshots
    for name, a, a32, a64 in zip(("k_lambda", "k_mu", "k_rho"), k, k32, k64):
        e, e32 = rel(a, a64), rel(a32, a64)
        print(f"{fmt} {name}: vs FP64 {e:.2e} (FP32 snapshots {e32:.2e})")
        assert e < tol


@pytest.mark.parametrize("fmt", ["f32", "f16", "bf16", "i16"])
@pytest.mark.parametrize("batch,sfe,ckpt", [(1, 10, 100), (3, 10, 50), (3, 1, 60), (2, 5, 300)])
def test_checkpointed_gradient_is_bitwise(gpu, fmt, batch, sfe, ckpt):
    P = inv.problems.tiny(sh=1, psv=1, nsrc=3, nt=300, sfe=sfe)
    obs = _obs(P)
    print(f"checkpoints {fmt} batch={batch} sfe={sfe} ckpt={ckpt}", flush=True)
    f0, k0 = _gradient(P, obs, batch, FMTS[fmt], 0, misfit=1, sigma=2)
    f1, k1 = _gradient(P, obs, batch, FMTS[fmt], ckpt, misfit=1, sigma=2)
    f2, k2 = _gradient(P, obs, batch, FMTS[fmt], ckpt, graphs=False, misfit=1, sigma=2)
    assert f0 == f1 == f2
    for a, b, c in zip(k0, k1, k2):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_checkpointed_adjoint_in_chunks(gpu):
    P = inv.problems.tiny(sh=1, psv=1, nsrc=2, nt=300, sfe=10)
    ctx = P.context(batch=2)
    ctx.set_snapshots(H.INV_SNAP_F16, 60)
    ctx.forward(0, 2, snapshots=True)
    rng = np.random.default_rng(5)
    for b in range(2):
        ctx.adstf_set(b, rng.standard_normal((P.ncomp, P.nrec, P.nt)).astype(np.float32))
    ctx.kernels_zero()
    ctx.adjoint(0, 2)
    whole = ctx.kernels_get()
    ctx.forward(0, 2, snapshots=True)
    ctx.kernels_zero()
    ctx.adjoint_begin(0, 2)
    pos = 0
    while pos < P.nt:
        pos = ctx.advance(37)
    chunked = ctx.kernels_get()
    ctx.close()
    for a, b in zip(whole, chunked):
        assert np.array_equal(a, b)


def test_unsupported_combinations_rejected(gpu):
    P = inv.problems.tiny(sh=0, psv=1, nsrc=1, nt=300, sfe=10)
    ctx = P.context(batch=1)
    with pytest.raises(H.InvError):
        ctx.set_snapshots(H.INV_SNAP_F16, 25)  # not a multiple of sfe
    with pytest.raises(H.InvError):
        ctx.set_snapshots(H.INV_SNAP_F32, 70)  # does not divide nt
    with pytest.raises(H.InvError):
        ctx.set_snapshots(7, 0)
    ctx.set_snapshots(H.INV_SNAP_BF16, 0)
    ctx.set_variant(H.INV_VARIANT_SPLIT)
    with pytest.raises(H.InvError):
        ctx.forward(0, 1, snapshots=True)
    ctx.set_variant(H.INV_VARIANT_FUSED)
    ctx.forward(0, 1, snapshots=True)
    ctx.close()


@pytest.mark.parametrize("fmt", ["f32", "f16", "bf16", "i16"])
@pytest.mark.parametrize("tile", [(4, 2), (16, 2), (4, 4), (8, 4)])
def test_wide_kernel_equals_column_kernel(gpu, fmt, tile):
    """The wide column kernel (2 or 4 columns per lane) gives bit for bit the
    one-column kernel's traces, misfit and FP32-snapshot gradient; with 16-bit
    snapshots (scaled per row segment, whose width is the tile's) and
    checkpoints the gradients agree to SEG_TOL."""
    P = inv.problems.tiny(sh=1, psv=1, nsrc=3, nt=300, sfe=10)
    print(f"wide tile {tile} {fmt}", flush=True)
    obs = _obs(P)
    out = []
    for variant in ("col", "wide"):
        ctx = P.context(batch=3)
        if variant == "wide":
            ctx.set_tile(*tile)
        ctx.set_snapshots(FMTS[fmt], 100)
        for s in range(P.nsrc):
            ctx.obs_set(s, obs[s])
        f = ctx.gradient(misfit_type=0, sigma=0)
        ctx.forward(0, 3, snapshots=False)
        tr = [ctx.traces_get(b) for b in range(3)]
        out.append((f, ctx.kernels_get(), tr))
        ctx.close()
    (fa, ka, ta), (fb, kb, tb) = out
    assert fa == fb
    for x, y in zip(ka, kb):
        if fmt == "f32":
            assert np.array_equal(x, y)
        else:
            e = rel(x, y)
            print(f"  {fmt} kernel difference between tiles: {e:.2e}")
            assert e <= SEG_TOL[fmt]
    for x, y in zip(ta, tb):
        assert np.array_equal(x, y)

"""inv_misfit_traces_noalloc (the waveform misfit of a batch of device traces
with no device allocation) against inv_misfit_trace, the existing per-trace
entry point, through the C ABI: adjoint sources and per-trace misfits equal
bit for bit, the total equals their FP64 sum; a second call with fewer traces
(the scratch keeps the first call's contents) gives its own results; more
traces than the scratch holds are refused."""
import numpy as np
import pytest

from conftest import inv

pytestmark = pytest.mark.gpu


def _traces(ctx, rng, ntr, nt):
    syn = rng.standard_normal((ntr, nt)).astype(np.float32)
    obs = (syn + 0.1 * rng.standard_normal((ntr, nt))).astype(np.float32)
    d = [ctx.alloc(ntr * nt * 4) for _ in range(3)]
    ctx.h2d(d[0], syn)
    ctx.h2d(d[1], obs)
    return d


def _reference(ctx, d, ntr, nt):
    """inv_misfit_trace trace by trace into a separate adjoint buffer."""
    ad = ctx.alloc(ntr * nt * 4)
    chi = np.array([ctx.misfit_trace(0, d[0] + 4 * r * nt, d[1] + 4 * r * nt, ad + 4 * r * nt)
                    for r in range(ntr)])
    out = ctx.d2h(ad, (ntr, nt))
    ctx.free(ad)
    return chi, out


def test_no_alloc_misfit_matches_per_trace_entry(gpu):
    P = inv.problems.tiny(nt=700, sh=0, psv=1)
    nt = P.nt
    print(f"no-alloc misfit: nt={nt}, 300 then 37 traces", flush=True)
    ctx = P.context(batch=1)
    rng = np.random.default_rng(11)
    for ntr in (300, 37):
        d = _traces(ctx, rng, ntr, nt)
        chi_ref, ad_ref = _reference(ctx, d, ntr, nt)
        chi, total = ctx.misfit_traces_noalloc(d[0], d[1], d[2], ntr)
        ad = ctx.d2h(d[2], (ntr, nt))
        assert np.array_equal(ad, ad_ref)
        assert np.array_equal(chi, chi_ref)
        assert abs(total - chi_ref.sum()) <= 1e-12 * abs(chi_ref.sum())
        for p in d:
            ctx.free(p)
    # more traces than the scratch holds: refused on the host (the buffers are
    # large enough that the call could not read or write out of them anyway)
    big = [ctx.alloc(65537 * nt * 4) for _ in range(3)]
    with pytest.raises(inv.invhip.InvError):
        ctx.misfit_traces_noalloc(big[0], big[1], big[2], 65537)
    for p in big:
        ctx.free(p)
    ctx.close()

"""The single-step hooks of the C ABI (inv_step_velocity, inv_step_stress,
inv_state_set): the reference's separable phases divS / divV
(solver/fdm.h:273-294) on the context's wavefields, against the oracle.

In exact math a velocity half-step followed by a stress half-step is one
time step of runForward, bit for bit: n of them from rest equal the oracle's
first n steps, and so does a run that starts with the fused kernels (whose
state alternates between two buffer sets) and continues with the hooks.
"""
import numpy as np
import pytest

from conftest import inv, oracle_problem

pytestmark = pytest.mark.gpu

CASES = [dict(sh=1, psv=0), dict(sh=0, psv=1), dict(sh=1, psv=1)]
IDS = ["sh", "psv", "sh+psv"]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_half_steps_from_rest_match_oracle(gpu, case):
    P = inv.problems.tiny(nt=120, **case)
    print(f"half-steps from rest {case}", flush=True)
    ctx = P.context(batch=1)
    ctx.set_math(inv.invhip.INV_MATH_EXACT)
    ctx.state_set(0, np.zeros((ctx.nstate, P.N), np.float32))
    O = oracle_problem(P)
    for it in range(40):
        ctx.step_velocity(1, 1, it)
        ctx.step_stress(1)
    _, _, ost = O.forward(1, nsteps=40)
    assert np.array_equal(ctx.state_get(0), ost)
    ctx.close()


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fused_run_continued_by_half_steps(gpu, case):
    """17 fused steps (odd: the state ends in the second buffer set), then 23
    half-step pairs; every source fires (first = -1)."""
    P = inv.problems.tiny(nt=120, **case)
    print(f"fused then half-steps {case}", flush=True)
    ctx = P.context(batch=1, adj=0)
    ctx.set_math(inv.invhip.INV_MATH_EXACT)
    ctx.set_tile(8, 1)
    ctx.set_graphs(False)
    ctx.set_step_limit(17)
    ctx.forward(-1, 1, snapshots=False)
    for it in range(17, 40):
        ctx.step_velocity(-1, 1, it)
        ctx.step_stress(1)
    _, _, ost = oracle_problem(P, adj=0).forward(-1, nsteps=40)
    assert np.array_equal(ctx.state_get(0), ost)
    ctx.close()


def test_state_set_round_trip_and_one_phase(gpu):
    """state_set then state_get returns the same fields; from a random
    stress field with v = u = 0, a stress half-step leaves the stress (v = 0
    has no gradient) and a velocity half-step with no force moves v by
    dt * divS / rho, which a second context computes the same way."""
    P = inv.problems.tiny(nt=60, sh=1, psv=1)
    ctx = P.context(batch=2)
    ctx.set_math(inv.invhip.INV_MATH_EXACT)
    rng = np.random.default_rng(3)
    st = rng.standard_normal((ctx.nstate, P.N)).astype(np.float32)
    ctx.state_set(1, st)
    assert np.array_equal(ctx.state_get(1), st)
    s0 = np.zeros_like(st)
    s0[2:4] = st[2:4]       # SH stress
    s0[8:11] = st[8:11]     # P-SV stress
    ctx.state_set(0, s0)
    ctx.step_stress(1)
    assert np.array_equal(ctx.state_get(0), s0)
    # a time step past the wavelets' support: no force
    P2 = P.with_(stf_x=np.zeros_like(P.stf_x), stf_y=np.zeros_like(P.stf_y),
                 stf_z=np.zeros_like(P.stf_z))
    ref = P2.context(batch=1)
    ref.set_math(inv.invhip.INV_MATH_EXACT)
    ref.state_set(0, s0)
    ref.step_velocity(0, 1, 0)
    ctx2 = P2.context(batch=1)
    ctx2.set_math(inv.invhip.INV_MATH_EXACT)
    ctx2.state_set(0, s0)
    ctx2.step_velocity(0, 1, 5)
    a, b = ref.state_get(0), ctx2.state_get(0)
    assert np.array_equal(a, b) and np.any(a[0]) and np.any(a[4])
    for c in (ctx, ref, ctx2):
        c.close()


def test_hooks_refuse_bad_arguments(gpu):
    P = inv.problems.tiny(nt=60, sh=0, psv=1)
    ctx = P.context(batch=1)
    with pytest.raises(inv.invhip.InvError):
        ctx.step_velocity(0, 1, 60)          # step out of range
    with pytest.raises(inv.invhip.InvError):
        ctx.step_velocity(0, 2, 0)           # more slots than the batch
    with pytest.raises(inv.invhip.InvError):
        ctx.state_set(1, np.zeros((ctx.nstate, P.N), np.float32))
    ctx.close()

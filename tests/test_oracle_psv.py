"""The oracle's P-SV path pinned to its SH path and to finite differences.

The SH forward of the oracle equals the reference's golden traces
(example/checker/traces, test_oracle_golden.py); the P-SV forward and the P-SV
kernels have no reference vectors.  Two facts of the scheme of
solver/fdm.h:76-219 tie them to SH:

* One-dimensional reductions.  With the model, the absorbing profile and the
  forcing uniform along one axis (every cell of a row or column driven by the
  same wavelet, the absorbing strips across that axis off), every difference
  along that axis is exactly zero (equal operands, or the out-of-range zero of
  divS*/divV*).  P-SV then splits into two SH problems:
    - force along the varying axis ("normal"): (v, u, normal stress) obey the
      SH update with mu' = lambda + 2 mu (updateSXZ, fdm.h:145-146), and the
      other normal stress is lambda / (lambda + 2 mu) times it;
    - force across it ("shear"): (v, u, sxz) obey the SH update with mu
      (fdm.h:147).
  The adjoint kernels (interaction*, fdm.h:202-219) reduce the same way:
  normal: k_lambda = k_mu' and k_mu = 2 k_mu' (J depends on lambda + 2 mu
  only); shear: k_mu = k_mu(SH), k_lambda = 0; k_rho = k_rho(SH) in both.
  Four reductions (fields uniform in x or in z) exercise every diagonal
  stiffness term, the lambda coupling terms of sxx / szz, both divergences and
  both directions of every stencil.  A stiffness mistake such as lambda + mu
  for lambda + 2 mu, a dropped coupling term or a swapped component fails
  them.
* Finite differences.  The kernels approximate the gradient of
  J = sum of squared tapered residuals (the functional whose derivative is the
  adjoint source 2 tw^2 r of misfit/waveform.h:6-7) divided by dt.  Away from
  the grid edges (where the zeroed derivatives of divS*/divV* are not the
  transpose of each other) and with the wavefield resolved, they agree with
  central differences to 1e-3; the reductions to 1e-4.  The 2-D check
  exercises the cross terms of the kernels that the reductions leave out
  (dux/dx duz/dz_fw in k_lambda, duz/dx dux/dz_fw in k_mu): dropping either
  moves its ratio to 0.30 / 0.86.
"""
import math

import numpy as np
import pytest

import pyoracle
from conftest import rel

F64 = np.float64


def _layers(n, seed, bump=False):
    """lambda, mu, rho per layer of 20 cells.  lambda and mu are integer
    multiples of 2^24 (below 2^20), so lambda + 2 mu is exact in FP32 and the
    SH problem with mu' = lambda + 2 mu sees the same stiffness."""
    rng = np.random.default_rng(seed)
    nl = (n + 19) // 20
    lam = np.repeat(rng.integers(300, 600, nl), 20)[:n].astype(F64)
    mu = np.repeat(rng.integers(200, 400, nl), 20)[:n].astype(F64)
    rho = np.repeat(rng.integers(1800, 2600, nl), 20)[:n].astype(F64)
    if bump:  # observed-data model: a smooth bump on the integer factors
        b = np.round(8 * np.exp(-((np.arange(n) - 90) ** 2) / 200.0))
        lam, mu, rho = lam + b, mu - b, rho + 10 * b
    s = 2.0 ** 24
    return (lam * s).astype(np.float32), (mu * s).astype(np.float32), rho.astype(np.float32)


def _reduction(axis, comp, bump=False, nu=6, nl=160, nt=600):
    """(P-SV problem, SH problem, P-SV component that moves) for fields
    uniform along `axis` ('x' or 'z'), forced along the varying axis
    (comp 'n') or across it ('s')."""
    lam1, mu1, rho1 = _layers(nl, 1, bump)
    if axis == "x":
        nx, nz = nu, nl
        spread = lambda a: np.tile(a, nx)  # [nx][nz], z fastest: same column everywhere
        ab = pyoracle.absbound(nx, nz, 20, 0.015, left=0, right=0, bottom=1, top=1)
        src = [(i, 40) for i in range(nx)]
        rec = [(i, j) for j in (30, 100) for i in range(nx)]
    else:
        nx, nz = nl, nu
        spread = lambda a: np.repeat(a, nz)
        ab = pyoracle.absbound(nx, nz, 20, 0.015, left=1, right=1, bottom=0, top=0)
        src = [(40, j) for j in range(nz)]
        rec = [(i, j) for i in (30, 100) for j in range(nz)]
    lam, mu, rho = spread(lam1), spread(mu1), spread(rho1)
    dt = np.float32(1e-3)
    w = pyoracle.ricker(nt, dt, 15.0, 0.08) * np.float32(1e9)
    zero = np.zeros(nt, np.float32)
    moving = "z" if (axis == "x") == (comp == "n") else "x"
    ns = len(src)
    rows = lambda a: np.stack([a] * ns)

    def mk(l, m, r, sh, psv, sx, sy, sz):
        return pyoracle.Problem(nx, nz, 10.0, 10.0, dt, nt, l, m, r, ab, [s[0] for s in src],
                                [s[1] for s in src], sx, sy, sz, [q[0] for q in rec],
                                [q[1] for q in rec], sh=sh, psv=psv, sfe=1, adj=1)

    P = mk(lam, mu, rho, 0, 1, rows(w if moving == "x" else zero), rows(zero),
           rows(w if moving == "z" else zero))
    m_sh = lam.astype(F64) + 2 * mu.astype(F64) if comp == "n" else mu.astype(F64)
    assert np.array_equal(m_sh.astype(np.float32).astype(F64), m_sh)
    S = mk(lam, m_sh.astype(np.float32), rho, 1, 0, rows(zero), rows(w), rows(zero))
    return P, S, moving


CASES = [("x", "n"), ("x", "s"), ("z", "n"), ("z", "s")]
IDS = ["x-uniform-normal", "x-uniform-shear", "z-uniform-normal", "z-uniform-shear"]


@pytest.mark.parametrize("axis,comp", CASES, ids=IDS)
def test_psv_forward_reduces_to_sh(axis, comp):
    P, S, mv = _reduction(axis, comp)
    tp, _, st = P.forward(-1, dtype=F64)
    ts, _, ss = S.forward(-1, dtype=F64)
    vx, vz, ux, uz, sxx, szz, sxz = st
    vy, uy, sxy, szy = ss
    still = "z" if mv == "x" else "x"
    assert np.abs(ts["y"]).max() > 1.0
    assert rel(tp[mv], ts["y"]) < 1e-13
    assert not np.any(tp[still])
    v, u, w = (vx, ux, None) if mv == "x" else (vz, uz, None)
    assert rel(v, vy) < 1e-13 and rel(u, uy) < 1e-13
    assert not np.any(vx if mv == "z" else vz) and not np.any(ux if mv == "z" else uz)
    # the SH stress along the varying axis: szy for x-uniform fields, sxy for z-uniform
    s_sh = szy if axis == "x" else sxy
    assert not np.any(sxy if axis == "x" else szy)
    lam, mu = P.lam.astype(F64), P.mu.astype(F64)
    if comp == "n":
        main, other = (szz, sxx) if axis == "x" else (sxx, szz)
        assert rel(main, s_sh) < 1e-13
        # the coupling term: lambda dvzdz in sxx (lambda dvxdx in szz)
        assert rel(other * (lam + 2 * mu), lam * main) < 1e-13
        assert not np.any(sxz)
    else:
        assert rel(sxz, s_sh) < 1e-13
        assert not np.any(sxx) and not np.any(szz)


def _adjoint_sources(Pr, obs):
    tr, snap, _ = Pr.forward(-1, dtype=F64, snapshots=True)
    comps = ["y"] if Pr.sh else ["x", "z"]
    ad = {c: np.stack([pyoracle.waveform(tr[c][r], obs[c][r], Pr.dt, F64)[1]
                       for r in range(Pr.nrec)]) for c in comps}
    return ad, snap


def _kernels(Pr, obs):
    ad, snap = _adjoint_sources(Pr, obs)
    kl, km, kr, _ = Pr.adjoint(ad, snap, dtype=F64)
    return kl, km, kr


def _taper(nt, dt):
    _, ad = pyoracle.waveform(np.ones(nt), np.zeros(nt), dt, dtype=F64)
    return np.sqrt(ad[::-1] / 2)


def _J(Pr, obs, isrc=-1):
    tr, _, _ = Pr.forward(isrc, dtype=F64)
    comps = ["y"] if Pr.sh else ["x", "z"]
    tw = _taper(Pr.nt, Pr.dt)
    return sum(float(np.sum(((tr[c] - obs[c]) * tw) ** 2)) for c in comps)


def _with_model(Pr, lam, mu, rho):
    q = lambda a: a.reshape(Pr.nsrc, Pr.nt)
    return pyoracle.Problem(Pr.nx, Pr.nz, Pr.dx, Pr.dz, Pr.dt, Pr.nt, lam, mu, rho, Pr.ab,
                            Pr.src_ix, Pr.src_iz, q(Pr.stf_x), q(Pr.stf_y), q(Pr.stf_z),
                            Pr.rec_ix, Pr.rec_iz, sh=Pr.sh, psv=Pr.psv, sfe=Pr.sfe, adj=1)


def _fd_ratio(Pr, obs, k, param, shape, J=_J):
    """fd * dt / (k . dm) for a 1e-3 relative perturbation of `param` with
    spatial shape `shape`; dm is the FP32-rounded difference actually
    applied."""
    arrs = [Pr.lam, Pr.mu, Pr.rho]
    base = arrs[param].astype(F64)
    d = shape * base * 1e-3
    plus, minus = list(arrs), list(arrs)
    plus[param] = (base + d).astype(np.float32)
    minus[param] = (base - d).astype(np.float32)
    dm = plus[param].astype(F64) - minus[param].astype(F64)
    fd = J(_with_model(Pr, *plus), obs) - J(_with_model(Pr, *minus), obs)
    return fd * float(Pr.dt) / float(np.dot(k, dm))


@pytest.mark.parametrize("axis,comp", CASES, ids=IDS)
def test_psv_kernels_reduce_to_sh(axis, comp):
    """k_lambda = k_mu', k_mu = 2 k_mu' (normal); k_mu = k_mu(SH), k_lambda = 0
    (shear); k_rho = k_rho(SH)."""
    Pt, St, _ = _reduction(axis, comp, bump=True)
    P, S, _ = _reduction(axis, comp)
    kp = _kernels(P, Pt.forward(-1, dtype=F64)[0])
    ks = _kernels(S, St.forward(-1, dtype=F64)[0])
    assert np.abs(ks[1]).max() > 0 and np.abs(ks[2]).max() > 0
    assert rel(kp[2], ks[2]) < 1e-12
    if comp == "n":
        assert rel(kp[0], ks[1]) < 1e-12
        assert rel(kp[1], 2 * ks[1]) < 1e-12
    else:
        assert not np.any(kp[0])
        assert rel(kp[1], ks[1]) < 1e-12


@pytest.mark.parametrize("axis,comp", CASES, ids=IDS)
def test_psv_reduction_kernels_match_finite_differences(axis, comp):
    """On the reductions the scheme is self-adjoint (no mixed derivatives) and
    the kernels equal central differences of J to 1e-4 (measured 3e-5)."""
    Pt, _, _ = _reduction(axis, comp, bump=True)
    P, _, _ = _reduction(axis, comp)
    obs = Pt.forward(-1, dtype=F64)[0]
    k = _kernels(P, obs)
    layer = np.exp(-((np.arange(160) - 70) ** 2) / 60.0)
    shape = np.tile(layer, P.nx) if axis == "x" else np.repeat(layer, P.nz)
    for param in ((0, 1, 2) if comp == "n" else (1, 2)):
        r = _fd_ratio(P, obs, k[param], param, shape)
        print(f"{axis}/{comp} param {param}: {r:.6f}")
        assert abs(r - 1) < 1e-4


def _resolved_psv():
    """60 x 60 cells of 10 m, 20-cell absorbing strips, homogeneous background
    (vp 3000, vs 1732, rho 2200) with a Gaussian +5 % anomaly for the observed
    data; two forces at 0.7 rad, nine receivers at 500 m depth; f0 10 Hz, so
    the shortest S wavelength spans about 7 cells."""
    n, dx, nt, dt = 60, 10.0, 500, np.float32(1e-3)
    i = np.arange(n)[:, None] * dx
    j = np.arange(n)[None, :] * dx

    def model(p):
        g = np.exp(-((i - 300) ** 2 + (j - 320) ** 2) / 60.0 ** 2)
        vp, vs, rho = 3000 * (1 + p * g), 1732 * (1 + p * g), 2200 * (1 + p * g)
        return ((rho * (vp ** 2 - 2 * vs ** 2)).astype(np.float32).ravel(),
                (rho * vs ** 2).astype(np.float32).ravel(), rho.astype(np.float32).ravel())

    ab = pyoracle.absbound(n, n, 20, 0.015)
    w = pyoracle.ricker(nt, dt, 10.0, 0.12) * np.float32(1e9)
    ca, sa = np.float32(math.cos(0.7)), np.float32(math.sin(0.7))
    src = [(15, 10), (45, 10)]
    rec = [(int(x / dx), 50) for x in np.linspace(100, 500, 9)]

    def prob(m):
        return pyoracle.Problem(n, n, dx, dx, dt, nt, *m, ab, [s[0] for s in src],
                                [s[1] for s in src], np.stack([w * ca] * 2),
                                np.stack([w] * 2), np.stack([w * sa] * 2),
                                [r[0] for r in rec], [r[1] for r in rec], sh=0, psv=1, sfe=1,
                                adj=1)

    return prob(model(0.0)), prob(model(0.05)), n, i, j


def test_psv_2d_kernels_match_finite_differences():
    """The full 2-D P-SV kernels, cross terms included, against central
    differences of J on a resolved problem away from the grid edges (measured
    within 1.2e-3)."""
    P, Pt, n, i, j = _resolved_psv()
    obs = [Pt.forward(s, dtype=F64)[0] for s in range(2)]
    ks = [np.zeros(n * n) for _ in range(3)]
    for s in range(2):
        tr, snap, _ = P.forward(s, dtype=F64, snapshots=True)
        ad = {c: np.stack([pyoracle.waveform(tr[c][r], obs[s][c][r], P.dt, F64)[1]
                           for r in range(P.nrec)]) for c in "xz"}
        P.adjoint(ad, snap, kernels=ks, dtype=F64)
    shape = np.exp(-((i - 300) ** 2 + (j - 300) ** 2) / 50.0 ** 2).ravel()
    J2 = lambda Q, o: sum(_J(Q, o[s], isrc=s) for s in range(2))
    for param in range(3):
        r = _fd_ratio(P, obs, ks[param], param, shape, J=J2)
        print(f"2-D param {param}: {r:.5f}")
        assert abs(r - 1) < 5e-3

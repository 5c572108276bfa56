"""The oracle against what the mathematics fixes (no reference run needed).

Each check is chosen so that a plausible transcription mistake fails it: a
flipped stencil offset breaks reciprocity, a wrong adjoint sign or a dropped
term breaks the finite-difference gradient check, a symmetric smoother fails
the one-sided brute force, a wrong Hilbert bin fails the numpy comparison.
"""
import numpy as np
import pytest

import pyoracle
from conftest import inv, oracle_problem, rel


def homogeneous(nx=60, nz=60, nt=220, sh=1, psv=0, ab=False, f0=30.0):
    f32 = np.float32
    vp, vs, rho = 3000.0, 1700.0, 2200.0
    lam = np.full(nx * nz, rho * (vp * vp - 2 * vs * vs), f32)
    mu = np.full(nx * nz, rho * vs * vs, f32)
    rh = np.full(nx * nz, rho, f32)
    abw = pyoracle.absbound(nx, nz, 10, 0.015) if ab else np.ones(nx * nz, f32)
    return lam, mu, rh, abw


def test_reciprocity_sh():
    """G(A->B) = G(B->A) for a y-force and vy receivers, homogeneous medium."""
    nx = nz = 60
    nt = 220
    lam, mu, rho, ab = homogeneous(nx, nz, nt)
    stf = pyoracle.source(nt, np.float32(1e-3), 30.0, 0.04, 0.0, 1e9)
    A, B = (22, 25), (35, 31)
    one = lambda a: np.stack([a])
    P1 = pyoracle.Problem(nx, nz, 10.0, 10.0, 1e-3, nt, lam, mu, rho, ab, [A[0]], [A[1]],
                          one(stf[0]), one(stf[1]), one(stf[2]), [B[0]], [B[1]], sh=1, psv=0,
                          adj=0)
    P2 = pyoracle.Problem(nx, nz, 10.0, 10.0, 1e-3, nt, lam, mu, rho, ab, [B[0]], [B[1]],
                          one(stf[0]), one(stf[1]), one(stf[2]), [A[0]], [A[1]], sh=1, psv=0,
                          adj=0)
    a, _, _ = P1.forward(0, dtype=np.float64)
    b, _, _ = P2.forward(0, dtype=np.float64)
    assert np.abs(a["y"]).max() > 0
    assert rel(a["y"], b["y"]) < 1e-10


def test_reciprocity_psv_same_component():
    """P-SV: x-force at A recorded as vx at B equals the swap.  The grid is
    large enough that no edge reflection (where the reference zeroes the
    derivatives) reaches the receivers within nt steps."""
    nx = nz = 130
    nt = 220
    lam, mu, rho, ab = homogeneous(nx, nz, nt)
    stf = pyoracle.source(nt, np.float32(1e-3), 30.0, 0.04, 0.0, 1e9)  # angle 0: x force
    A, B = (58, 60), (72, 68)
    one = lambda a: np.stack([a])
    mk = lambda s, r: pyoracle.Problem(nx, nz, 10.0, 10.0, 1e-3, nt, lam, mu, rho, ab, [s[0]],
                                       [s[1]], one(stf[0]), one(stf[1]), one(stf[2]), [r[0]],
                                       [r[1]], sh=0, psv=1, adj=0)
    a, _, _ = mk(A, B).forward(0, dtype=np.float64)
    b, _, _ = mk(B, A).forward(0, dtype=np.float64)
    assert np.abs(a["x"]).max() > 0
    assert rel(a["x"], b["x"]) < 1e-10


def test_no_growth_without_absorption():
    """ab = 1, source finished early: velocity energy stays bounded (no blow-up)
    and the wave reaches the receivers."""
    nx = nz = 50
    nt = 600
    lam, mu, rho, ab = homogeneous(nx, nz, nt)
    stf = pyoracle.source(nt, np.float32(1e-3), 40.0, 0.03, 0.5, 1e9)
    one = lambda a: np.stack([a])
    P = pyoracle.Problem(nx, nz, 10.0, 10.0, 1e-3, nt, lam, mu, rho, ab, [25], [25],
                         one(stf[0]), one(stf[1]), one(stf[2]), [10, 40], [25, 25], sh=1, psv=1,
                         adj=0)
    e = []
    for n in (150, 300, 450, 600):
        _, _, st = P.forward(0, nsteps=n, dtype=np.float64)
        e.append(float(np.sum(st[0] ** 2) + np.sum(st[4] ** 2) + np.sum(st[5] ** 2)))
    assert all(np.isfinite(e)) and e[0] > 0
    assert max(e[1:]) < 3.0 * e[0]


def test_fp32_tracks_fp64():
    P = inv.problems.tiny(sh=1, psv=1)
    O = oracle_problem(P)
    for isrc in range(P.nsrc):
        a, _, sa = O.forward(isrc)
        b, _, sb = O.forward(isrc, dtype=np.float64)
        for c in "xyz":
            assert rel(a[c], b[c]) < 1e-5
        assert rel(sa, sb) < 1e-5


def _taper(nt, dt):
    ones = np.ones(nt)
    _, ad = pyoracle.waveform(ones, np.zeros(nt), dt, dtype=np.float64)
    return np.sqrt(ad[::-1] / 2)


def _kernels(P, obs, dtype=np.float64):
    O = oracle_problem(P)
    ks = [np.zeros(P.N, dtype) for _ in range(3)]
    comps = (["y"] if P.sh else []) + (["x", "z"] if P.psv else [])
    for s in range(P.nsrc):
        tr, snap, _ = O.forward(s, dtype=dtype, snapshots=True)
        ad = {c: np.stack([pyoracle.waveform(tr[c][r], obs[s][c][r], P.dt, dtype)[1]
                           for r in range(P.nrec)]) for c in comps}
        O.adjoint(ad, snap, kernels=ks, dtype=dtype)
    return ks


def _resolved(sh, psv, sfe):
    return inv.problems.tiny(sh=sh, psv=psv, nx=60, nz=50, nt=500, nsrc=2, nrec=12, sfe=sfe,
                             f0=8.0, t0=0.15)


@pytest.mark.parametrize("case,param", [((1, 0), "mu"), ((1, 0), "rho")])
def test_adjoint_kernel_sh_finite_differences(case, param):
    """The reference's SH kernels (solver/fdm.h:198-210) equal the gradient of
    J = sum of squared tapered residuals (the functional of the adjoint source
    2 tw^2 r, misfit/waveform.h:6-7) divided by dt, checked by central
    differences of a smooth 0.1 % blob perturbation (measured ratios 1.000 and
    1.002).  The P-SV kernels are checked the same way in test_oracle_psv.py,
    on a problem whose receivers sit away from the grid edges: this one
    records at row 4 and has 8-cell absorbing strips, so waves reach the grid
    edges, where the zeroed derivatives of divS*/divV* are not each other's
    transpose; P-SV's mode conversions make that visible (ratios 0.56-1.54
    here), SH does not."""
    P = _resolved(*case, sfe=1)
    Ot = oracle_problem(P, model="true")
    obs = [Ot.forward(s, dtype=np.float64)[0] for s in range(P.nsrc)]
    ks = _kernels(P, obs)
    k = {"lambda": ks[0], "mu": ks[1], "rho": ks[2]}[param]
    key = {"lambda": "lam0", "mu": "mu0", "rho": "rho0"}[param]
    base = getattr(P, key).astype(np.float64)
    i = np.arange(P.nx)[:, None]
    j = np.arange(P.nz)[None, :]
    dm = (np.exp(-((i - 30) ** 2 + (j - 25) ** 2) / 50.0).ravel() * base * 1e-3)
    comps = (["y"] if P.sh else []) + (["x", "z"] if P.psv else [])

    def J(m):
        O = oracle_problem(P.with_(**{key: m.astype(np.float32)}))
        tw = _taper(P.nt, P.dt)
        tot = 0.0
        for s in range(P.nsrc):
            tr, _, _ = O.forward(s, dtype=np.float64)
            for c in comps:
                tot += float(np.sum(((tr[c] - obs[s][c]) * tw) ** 2))
        return tot

    fd = (J(base + dm) - J(base - dm)) / 2
    ratio = fd * float(P.dt) / float(np.dot(k, dm))
    print(f"{param}: fd*dt / (k . dm) = {ratio:.4f}")
    assert abs(ratio - 1) < 0.01


@pytest.mark.parametrize("case", [(0, 1), (1, 0)], ids=["psv", "sh"])
def test_subsampled_correlation_matches_full(case):
    """Correlating every sfe = 10 steps with weight sfe*dt (fdm.h:583-585)
    approximates the every-step sum: pins the snapshot pairing (adjoint step
    it with forward step nt-1-it) and the ndt scale."""
    P1 = _resolved(*case, sfe=1)
    Ot = oracle_problem(P1, model="true")
    obs = [Ot.forward(s, dtype=np.float64)[0] for s in range(P1.nsrc)]
    k1 = _kernels(P1, obs)
    k10 = _kernels(P1.with_(sfe=10), obs)
    for a, b in zip(k1, k10):
        if np.any(a):
            e = rel(b, a)
            print(f"sfe 10 vs 1: {e:.3e}")
            assert e < 0.02


def test_gauss_one_sided_brute_force():
    """filter/gaussian.h with the size_t wrap: only n <= i weights, normalised
    by the product of the one-sided weight sums."""
    rng = np.random.default_rng(5)
    nx, nz, sigma = 23, 17, 3
    a = rng.standard_normal((nx, nz))
    w = lambda m: np.exp(-m * m / (2.0 * sigma * sigma)) / (np.sqrt(2 * np.pi) * sigma)
    out = np.zeros_like(a)
    for i in range(nx):
        for j in range(nz):
            num = sum(w(i - n) * w(j - m) * a[n, m] for n in range(i + 1) for m in range(j + 1))
            den = sum(w(i - n) for n in range(i + 1)) * sum(w(j - m) for m in range(j + 1))
            out[i, j] = num / den
    got = pyoracle.gauss(a.ravel(), nx, nz, sigma, dtype=np.float64).reshape(nx, nz)
    assert rel(got, out) < 1e-6
    # a symmetric smoother differs clearly
    assert not np.allclose(got[0], a[0])  # row 0 mixes only z-neighbours before it
    assert np.allclose(got[0, 0], a[0, 0])


def test_waveform_closed_form():
    nt, dt = 500, np.float32(2e-3)
    syn = np.zeros(nt)
    obs = np.zeros(nt)
    chi, ad = pyoracle.waveform(syn, obs, dt, dtype=np.float64)
    assert chi == 0 and not np.any(ad)
    c = 0.75
    chi, ad = pyoracle.waveform(syn + c, obs, dt, dtype=np.float64)
    tw = _taper(nt, dt)
    assert tw[0] == 0 and tw[nt // 2] == 1
    assert abs(chi - c * np.sqrt(np.sum(tw * tw)) * np.sqrt(float(dt))) < 1e-12
    assert np.allclose(ad[::-1], 2 * c * tw * tw, rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [16, 17, 30, 45])
def test_hilbert_mask_matches_numpy(n):
    """Analytic signal IFFT(mask * FFT(x)), unnormalised, with bin n/2+1 left
    untouched for even n (misfit/envelope.h:4-28)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    X = np.fft.fft(x)
    m = np.zeros(n)
    m[0] = 1
    for i in range(1, n):
        if n % 2 == 0:
            m[i] = 2 if i < n // 2 + 1 else (1 if i == n // 2 + 1 else 0)
        else:
            m[i] = 2 if i < (n + 1) // 2 else 0
    ref = np.fft.ifft(X * m) * n
    got = pyoracle.hilbert(x)
    assert np.allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


def _envelope_case(n, k, a_syn, a_obs):
    t = np.arange(n)
    th = 2 * np.pi * k * t / n
    return np.cos(th) * a_syn, np.cos(th) * a_obs, th


@pytest.mark.parametrize("n,k", [(400, 7), (375, 11), (64, 5), (45, 4)])
@pytest.mark.parametrize("a_syn,a_obs", [(3.0, 2.0), (0.4, 0.9)], ids=["floor-rel", "floor-abs"])
def test_envelope_closed_form(n, k, a_syn, a_obs):
    """EnvelopeMisfit::run (misfit/envelope.h:62-86) on syn = A cos(2 pi k t / n),
    obs = B cos(2 pi k t / n), 0 < k < n/2 - 1: the analytic signal is
    n A e^{i theta}, so E_syn = |A|, E_obs = |B| and H[syn] = A sin(theta);
    etmp = c = (|A| - |B|) / (|A| + floor) with floor = max(0.05 |A|, 0.05)
    (both branches: |A| = 3 and 0.4); H[c A sin] = -c A cos, so the adjoint
    source is adstf[n-1-t] = 2 c A cos(theta_t) (time-reversed) and
    chi = ||A| - |B|| sqrt(n dt).  Even and odd n."""
    dt = 1e-3
    syn, obs, th = _envelope_case(n, k, a_syn, a_obs)
    chi, ad = pyoracle.envelope(syn, obs, dt, dtype=np.float64)
    floor = max(0.05 * abs(a_syn), 0.05)
    c = (abs(a_syn) - abs(a_obs)) / (abs(a_syn) + floor)
    want = 2 * c * a_syn * np.cos(th)
    assert np.allclose(ad[::-1], want, rtol=0, atol=1e-11)
    assert abs(chi - abs(abs(a_syn) - abs(a_obs)) * np.sqrt(n * dt)) < 1e-11
    # a floor of 0.05 |A| where 0.05 applies (or the reverse) changes c
    other = 0.05 if floor != 0.05 else 0.05 * abs(a_syn)
    c2 = (abs(a_syn) - abs(a_obs)) / (abs(a_syn) + other)
    assert not np.allclose(ad[::-1], 2 * c2 * a_syn * np.cos(th), rtol=0, atol=1e-6)


def test_envelope_bin_quirk_closed_form():
    """Even n with k = n/2 - 1: the mirror bin n - k = n/2 + 1 is the one the
    mask leaves untouched (misfit/envelope.h:11-16), so the analytic signal
    of cos(theta) is n e^{i theta} + (n/2) e^{-i theta} and the envelope is
    sqrt(5/4 + cos(2 theta)) instead of 1."""
    n = 64
    k = n // 2 - 1
    syn, obs, th = _envelope_case(n, k, 1.0, 0.0)
    a = pyoracle.hilbert(syn)
    want = n * np.exp(1j * th) + n / 2 * np.exp(-1j * th)
    assert np.allclose(a, want, rtol=0, atol=1e-9)
    # the misfit sees that envelope: chi = ||E_syn|| sqrt(dt) with obs = 0
    dt = 1e-3
    chi, _ = pyoracle.envelope(syn, obs, dt, dtype=np.float64)
    assert abs(chi - np.linalg.norm(np.sqrt(1.25 + np.cos(2 * th))) * np.sqrt(dt)) < 1e-10


def test_ricker_closed_form():
    nt, dt, f0, t0 = 400, np.float32(1e-3), 12.0, 0.1
    stf = pyoracle.ricker(nt, dt, f0, t0).astype(np.float64)
    t = np.arange(nt) * float(dt) - t0
    ref = -t * np.exp(-(np.pi * f0 * t) ** 2)
    ref /= np.abs(ref).max()
    assert np.allclose(stf, ref, atol=2e-6)


def test_geometry_helpers():
    assert pyoracle.grid_dims(40401, 480000.0, 480000.0)[:2] == (201, 201)
    assert pyoracle.grid_dims(461 * 151, 9200.0, 3000.0)[:2] == (461, 151)
    assert pyoracle.calc_idx(0.0, 100.0, 11) == 0
    assert pyoracle.calc_idx(100.0, 100.0, 11) == 10
    assert pyoracle.calc_idx(54.9, 100.0, 11) == 5


def _explosive(nx, nz, nt, src, stf_y, ab, src_type=1, stf_xz=None):
    lam, mu, rho, _ = homogeneous(nx, nz, nt)
    one = lambda a: np.stack([np.asarray(a, np.float32)])
    sx, sz = (np.zeros(nt, np.float32),) * 2 if stf_xz is None else stf_xz
    return pyoracle.Problem(nx, nz, 10.0, 10.0, 1e-3, nt, lam, mu, rho, ab, [src[0]], [src[1]],
                            one(sx), one(stf_y), one(sz), [2], [2], sh=0, psv=1, adj=0,
                            src_type=[src_type])


def test_explosive_source_first_step():
    """Explosive source (source type 1, no reference counterpart): from rest,
    one step leaves exactly stf_y[0] in sxx and szz at the source cell and
    nothing anywhere else.  The point-force wavelets stf_x / stf_z are set
    and must not be injected."""
    nx, nz, nt = 30, 26, 4
    src = (12, 17)
    stf_y = np.array([2.5, 0.0, 0.0, 0.0], np.float32)
    junk = (np.full(nt, 7.0, np.float32), np.full(nt, -3.0, np.float32))
    P = _explosive(nx, nz, nt, src, stf_y, np.ones(nx * nz, np.float32), stf_xz=junk)
    _, _, st = P.forward(0, nsteps=1, dtype=np.float64)
    vx, vz, ux, uz, sxx, szz, sxz = st
    k = src[0] * nz + src[1]
    want = np.zeros(nx * nz)
    want[k] = 2.5
    assert np.array_equal(sxx, want) and np.array_equal(szz, want)
    for f in (vx, vz, ux, uz, sxz):
        assert not np.any(f)
    # the same wavelets as a point force move the velocity instead
    F = _explosive(nx, nz, nt, src, stf_y, np.ones(nx * nz, np.float32), src_type=0,
                   stf_xz=junk)
    _, _, st = F.forward(0, nsteps=1, dtype=np.float64)
    assert st[0][k] != 0 and st[1][k] != 0


def test_explosive_source_xz_symmetry():
    """The scheme is symmetric under exchanging x and z on a square
    homogeneous grid with all four absorbing strips (the update equations map
    onto each other, vx <-> vz, sxx <-> szz), so an explosion at the centre
    gives vx(i, j) = vz(j, i) and sxx(i, j) = szz(j, i); an injection into one
    component only, a sign error or an off-centre cell breaks it.  A point
    force (the control) does not have the symmetry."""
    n, nt = 41, 150
    ab = pyoracle.absbound(n, n, 10, 0.015)
    stf = pyoracle.source(nt, np.float32(1e-3), 30.0, 0.04, 0.0, 1e9)
    P = _explosive(n, n, nt, (20, 20), stf[1], ab)
    _, _, st = P.forward(0, dtype=np.float64)
    vx, vz, ux, uz, sxx, szz, sxz = (f.reshape(n, n) for f in st)
    assert np.abs(vx).max() > 0
    assert rel(vx, vz.T) < 1e-10 and rel(ux, uz.T) < 1e-10
    assert rel(sxx, szz.T) < 1e-10 and rel(sxz, sxz.T) < 1e-10
    F = _explosive(n, n, nt, (20, 20), stf[1], ab, src_type=0, stf_xz=(stf[0], stf[2]))
    _, _, st = F.forward(0, dtype=np.float64)
    assert rel(st[0].reshape(n, n), st[1].reshape(n, n).T) > 0.1

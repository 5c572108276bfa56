"""Synthetic inversion setups (model, acquisition, time stepping).

Each generator returns a `Problem`: true and initial models on the grid
(lambda, mu, rho, z fastest), point sources with Ricker wavelets from the
library's own RickerSource (inv_source_ricker), receivers, and the solver
flags.  `checker` is the adjacent-squares checkerboard of the 200x200
gradient workload; `checker_gapped` follows the reference's
util/generate_checker.cu:4-70 (squares separated by gaps, optional one-sided
smoothing) and reproduces its example model; the layered/faulted generator
is a Marmousi-shaped model of our own.  `write_config` writes any of them as a
configuration directory for the command-line driver.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from . import invhip


def vps2lm(vp, vs, rho):
    """vps2lm of solver/fdm.h:25-36 in FP32 (no fused multiply-add)."""
    vp = np.asarray(vp, np.float32)
    vs = np.asarray(vs, np.float32)
    rho = np.asarray(rho, np.float32)
    lam = np.where(vp > vs, rho * (vp * vp - np.float32(2) * vs * vs), np.float32(0))
    mu = rho * vs * vs
    return lam.astype(np.float32), mu.astype(np.float32)


@dataclass
class Problem:
    name: str
    nx: int
    nz: int
    dx: float
    dz: float
    dt: float
    nt: int
    sh: int
    psv: int
    lam: np.ndarray          # true model
    mu: np.ndarray
    rho: np.ndarray
    lam0: np.ndarray         # initial model
    mu0: np.ndarray
    rho0: np.ndarray
    src_ix: np.ndarray
    src_iz: np.ndarray
    stf_x: np.ndarray        # [nsrc][nt]
    stf_y: np.ndarray
    stf_z: np.ndarray
    rec_ix: np.ndarray
    rec_iz: np.ndarray
    sfe: int = 10
    abs_width: int = 20
    abs_alpha: float = 0.015
    abs_left: int = 1
    abs_right: int = 1
    abs_top: int = 1
    abs_bottom: int = 1
    inv_lambda: int = 1
    inv_mu: int = 1
    inv_rho: int = 1
    trace_type: int = 0
    misfit: int = 0
    sigma: int = 0
    src_type: np.ndarray = None  # per source: 0 point force (default), 1 explosive
    meta: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.nx * self.nz

    @property
    def nsrc(self):
        return len(self.src_ix)

    @property
    def nrec(self):
        return len(self.rec_ix)

    @property
    def ncomp(self):
        return (1 if self.sh else 0) + (2 if self.psv else 0)

    def cell_updates(self, nshots=None, adjoint=True):
        """Cell-updates of one gradient: sum over shots of nx*nz*nt*(2 with adjoint)."""
        ns = self.nsrc if nshots is None else nshots
        return ns * self.nx * self.nz * self.nt * (2 if adjoint else 1)

    def grid_kwargs(self, batch=1, adj=1):
        return dict(nx=self.nx, nz=self.nz, dx=self.dx, dz=self.dz, dt=self.dt, nt=self.nt,
                    sh=self.sh, psv=self.psv, adj=adj, trace_type=self.trace_type, sfe=self.sfe,
                    abs_width=self.abs_width, abs_alpha=self.abs_alpha, abs_left=self.abs_left,
                    abs_right=self.abs_right, abs_top=self.abs_top,
                    abs_bottom=self.abs_bottom, inv_lambda=self.inv_lambda, inv_mu=self.inv_mu,
                    inv_rho=self.inv_rho, batch=batch)

    def context(self, batch=1, adj=1, model="init", device=0):
        """A library context with grid, model and acquisition set."""
        ctx = invhip.Context(device)
        ctx.grid_set(**self.grid_kwargs(batch=batch, adj=adj))
        self.load_model(ctx, model)
        ctx.sources_set(self.src_ix, self.src_iz, self.stf_x, self.stf_y, self.stf_z)
        if self.src_type is not None and np.any(self.src_type):
            ctx.sources_type(self.src_type)
        ctx.receivers_set(self.rec_ix, self.rec_iz)
        return ctx

    def load_model(self, ctx, model="init"):
        if model == "true":
            ctx.model_upload(self.lam, self.mu, self.rho)
        else:
            ctx.model_upload(self.lam0, self.mu0, self.rho0)

    def with_(self, **kw):
        return replace(self, **kw)


def _wavelets(nt, dt, f0, t0, amp, angle, nsrc):
    sx, sy, sz = invhip.source_ricker(nt, dt, f0, t0, angle, amp)
    rep = lambda a: np.repeat(a[None, :], nsrc, axis=0).astype(np.float32)
    return rep(sx), rep(sy), rep(sz)


def _smooth(a, sigma):
    """Symmetric separable Gaussian smoothing, radius 3 sigma, edges mirrored
    (a[-k] = a[k]), float64."""
    if sigma <= 0:
        return a.astype(np.float64)
    from scipy.ndimage import gaussian_filter1d

    out = a.astype(np.float64)
    for ax in (0, 1):
        out = gaussian_filter1d(out, sigma, axis=ax, mode="mirror", truncate=3.0)
    return out


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------
def checker(nx=200, nz=200, dx=10.0, dt=1e-3, nt=2000, cell=20, frac=0.10, vp0=3000.0,
            vs0=1732.0, rho0=2200.0, nsrc=4, nrec=40, f0=10.0, t0=0.15, amp=1e10, sh=0,
            psv=1, src_depth=20, rec_depth=10, sigma=5, explosive=False, name="checker"):
    """P-SV checkerboard gradient setup: background +-frac on vp, vs, rho in
    squares of `cell` cells, sign (-1)^(floor(i/cell)+floor(j/cell)); the
    initial model is the background.  Sources evenly spaced in x at
    `src_depth` cells (explosive ones with explosive=True), receivers at
    `rec_depth` cells."""
    i = np.arange(nx)[:, None]
    j = np.arange(nz)[None, :]
    sign = np.where(((i // cell) + (j // cell)) % 2 == 0, 1.0, -1.0)
    f32 = np.float32
    vp = (vp0 * (1 + frac * sign)).astype(f32)
    vs = (vs0 * (1 + frac * sign)).astype(f32)
    rho = (rho0 * (1 + frac * sign)).astype(f32)
    lam, mu = vps2lm(vp.ravel(), vs.ravel(), rho.ravel())
    b = np.ones(nx * nz, f32)
    lam0, mu0 = vps2lm(b * f32(vp0), b * f32(vs0), b * f32(rho0))
    xmax = (nx - 1) * dx
    src_x = [(s + 1) * xmax / (nsrc + 1) for s in range(nsrc)]
    rec_x = [(r + 0.5) * xmax / nrec for r in range(nrec)]
    src_ix = np.array([invhip.calc_idx(x, xmax, nx) for x in src_x], np.int32)
    rec_ix = np.array([invhip.calc_idx(x, xmax, nx) for x in rec_x], np.int32)
    sx, sy, sz = _wavelets(nt, dt, f0, t0, amp, 0.0, nsrc)
    return Problem(name, nx, nz, f32(dx), f32(dx), f32(dt), nt, sh, psv, lam, mu,
                   rho.ravel(), lam0, mu0, b * f32(rho0), src_ix,
                   np.full(nsrc, src_depth, np.int32), sx, sy, sz, rec_ix,
                   np.full(nrec, rec_depth, np.int32), sigma=sigma,
                   src_type=np.full(nsrc, 1 if explosive else 0, np.int32),
                   meta=dict(f0=f0, t0=t0, amp=amp))


def layered_faulted(nx=737, nz=240, dx=12.5, seed=1, nwater=12, vmin=1500.0, vmax=4700.0,
                    nfaults=3):
    """Marmousi-shaped vp/vs/rho: vp rising with depth below a water layer,
    8-15 seeded dipping layers, `nfaults` normal faults with 50-200 m throw;
    vs = vp/1.73 and Gardner density 310 vp^0.25 below the water."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None] * dx
    z = np.arange(nz)[None, :] * dx
    xmax, zmax = (nx - 1) * dx, (nz - 1) * dx
    zw = nwater * dx
    zeff = np.broadcast_to(z, (nx, nz)).astype(np.float64).copy()
    for _ in range(nfaults):
        x0 = rng.uniform(0.15, 0.85) * xmax
        dip = np.deg2rad(rng.uniform(55.0, 75.0))
        throw = rng.uniform(50.0, 200.0)
        hanging = x > x0 + z / np.tan(dip)
        zeff = np.where(hanging, zeff - throw, zeff)
    nl = int(rng.integers(8, 16))
    bases = np.sort(rng.uniform(zw, zmax, nl))
    slopes = rng.uniform(-0.12, 0.12, nl)
    layer = np.zeros((nx, nz), np.int64)
    for b0, s in zip(bases, slopes):
        layer += (zeff > b0 + s * (x - xmax / 2)).astype(np.int64)
    trend = vmin + (vmax - vmin) * np.clip((zeff - zw) / (zmax - zw), 0, 1)
    jitter = rng.uniform(-250.0, 250.0, nl + 1)
    vp = np.clip(trend + jitter[layer], vmin, vmax)
    vs = vp / 1.73
    rho = 310.0 * vp ** 0.25
    water = np.broadcast_to(np.arange(nz)[None, :] < nwater, (nx, nz))
    vp = np.where(water, 1500.0, vp)
    vs = np.where(water, 0.0, vs)
    rho = np.where(water, 1000.0, rho)
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32), water


def marmousi(nx=737, nz=240, dx=12.5, dt=1e-3, nt=4000, nsrc=16, nrec=368, f0=5.0, t0=0.3,
             amp=1e10, seed=1, src_depth=2, rec_depth=2, rec_step=2, smooth=40.0, sh=0, psv=1,
             name="marmousi", sigma=0):
    """Batched P-SV inversion setup on the Marmousi-shaped model.  The
    initial model is the true one smoothed with a symmetric Gaussian
    (`smooth` cells), water rows kept.

    Receivers sit on row 2, the shallowest row the stencil couples to the
    interior: the z-terms of the divergence exist only for 2 <= j < nz-2
    (solver/fdm.h:84,98) and those of the gradient for 1 <= j < nz-2
    (fdm.h:111,128), so rows 0 and 1 never receive energy from below and a
    receiver there records zeros."""
    vp, vs, rho, water = layered_faulted(nx, nz, dx, seed)
    f32 = np.float32
    sm = lambda a: np.where(water, a, _smooth(a, smooth)).astype(f32)
    vp0, vs0, rho0 = sm(vp), sm(vs), sm(rho)
    vs0 = np.where(water, f32(0), vs0).astype(f32)
    lam, mu = vps2lm(vp.ravel(), vs.ravel(), rho.ravel())
    lam0, mu0 = vps2lm(vp0.ravel(), vs0.ravel(), rho0.ravel())
    xmax = (nx - 1) * dx
    src_x = [(s + 0.5) * xmax / nsrc for s in range(nsrc)]
    src_ix = np.array([invhip.calc_idx(xx, xmax, nx) for xx in src_x], np.int32)
    rec_ix = (np.arange(nrec) * rec_step).astype(np.int32)
    rec_ix = np.minimum(rec_ix, nx - 1).astype(np.int32)
    sx, sy, sz = _wavelets(nt, dt, f0, t0, amp, 0.0, nsrc)
    return Problem(name, nx, nz, f32(dx), f32(dx), f32(dt), nt, sh, psv, lam, mu, rho.ravel(),
                   lam0, mu0, rho0.ravel(), src_ix, np.full(nsrc, src_depth, np.int32), sx, sy,
                   sz, rec_ix, np.full(nrec, rec_depth, np.int32), abs_top=0, sigma=sigma,
                   meta=dict(f0=f0, t0=t0, amp=amp, seed=seed))


def tiny(nx=48, nz=40, dx=10.0, dt=1e-3, nt=300, nsrc=3, nrec=12, sh=1, psv=1, seed=7,
         f0=25.0, t0=0.04, sfe=10, abs_width=8, name="tiny"):
    """Small random layered setup for parity tests (seconds on the oracle)."""
    rng = np.random.default_rng(seed)
    f32 = np.float32
    z = np.arange(nz)[None, :]
    base = 2000.0 + 1500.0 * z / (nz - 1) + np.zeros((nx, 1))
    vp = (base * (1 + 0.05 * rng.standard_normal((nx, nz)))).astype(f32)
    vs = (vp / 1.8).astype(f32)
    rho = (1800.0 + 0.2 * vp).astype(f32)
    lam, mu = vps2lm(vp.ravel(), vs.ravel(), rho.ravel())
    vp0 = _smooth(vp, 3.0).astype(f32)
    vs0 = (vp0 / 1.8).astype(f32)
    rho0 = (1800.0 + 0.2 * vp0).astype(f32)
    lam0, mu0 = vps2lm(vp0.ravel(), vs0.ravel(), rho0.ravel())
    src_ix = rng.integers(abs_width, nx - abs_width, nsrc).astype(np.int32)
    src_iz = rng.integers(3, nz // 3, nsrc).astype(np.int32)
    rec_ix = np.linspace(2, nx - 3, nrec).astype(np.int32)
    rec_iz = np.full(nrec, 4, np.int32)
    angle = 0.7
    stf = [invhip.source_ricker(nt, dt, f0, t0, angle, 1e9) for _ in range(nsrc)]
    sx = np.stack([s[0] for s in stf])
    sy = np.stack([s[1] for s in stf])
    sz = np.stack([s[2] for s in stf])
    return Problem(name, nx, nz, f32(dx), f32(dx), f32(dt), nt, sh, psv, lam, mu, rho.ravel(),
                   lam0, mu0, rho0.ravel(), src_ix, src_iz, sx, sy, sz, rec_ix, rec_iz, sfe=sfe,
                   abs_width=abs_width, meta=dict(f0=f0, t0=t0, seed=seed))


# ---------------------------------------------------------------------------
# mixed-size shot-solve requests for the request server
# ---------------------------------------------------------------------------
SERVE_GRIDS = ((256, 256), (512, 512), (1024, 512), (2048, 1024))


def layered_random(nx, nz, seed, vmin=1500.0, vmax=4700.0):
    """Seeded random layered model: 3-8 flat layers whose vp rises with depth
    between vmin and vmax (+-150 m/s jitter), vs = vp/1.73, Gardner density
    310 vp^0.25.  Returns lambda, mu, rho (FP32, z fastest)."""
    rng = np.random.default_rng(seed)
    nl = int(rng.integers(3, 9))
    tops = np.sort(rng.integers(1, nz - 1, nl - 1))
    vp_l = np.sort(rng.uniform(vmin, vmax, nl)) + rng.uniform(-150.0, 150.0, nl)
    vp_l = np.clip(vp_l, vmin, vmax)
    layer = np.searchsorted(tops, np.arange(nz), side="right")
    vp = np.broadcast_to(vp_l[layer][None, :], (nx, nz)).astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    lam, mu = vps2lm(vp.ravel(), vs.ravel(), rho.ravel())
    return lam, mu, rho.ravel()


def serving_request(rid, seed=3, dx=10.0, dt=1e-3, sfe=10, f0=8.0, t0=0.15, amp=1e10,
                    grids=SERVE_GRIDS, nt_range=(1000, 6000), max_shots=8):
    """Request `rid` of the mixed-size serving load: grid drawn uniformly from
    `grids`, nt uniform over the multiples of sfe in nt_range, SH or P-SV and
    forward-only or forward+adjoint with probability 1/2 each, 1..max_shots
    shots evenly spaced at 2 cells depth, a random layered model seeded
    1000 + rid, receivers every 4 cells on row 2 (the same for every request
    on a grid, so like-sized requests can share a launch).  Observed traces of
    adjoint requests are zero (the misfit is then the tapered trace energy;
    the work is the same as for real data)."""
    rng = np.random.default_rng([seed, rid])
    nx, nz = grids[int(rng.integers(len(grids)))]
    nt = int(rng.integers(nt_range[0] // sfe, nt_range[1] // sfe + 1)) * sfe
    sh = bool(rng.random() < 0.5)
    adjoint = bool(rng.random() < 0.5)
    ns = int(rng.integers(1, max_shots + 1))
    xmax = (nx - 1) * dx
    src_ix = np.array([invhip.calc_idx((s + 0.5) * xmax / ns, xmax, nx) for s in range(ns)],
                      np.int32)
    sx, sy, sz = _wavelets(nt, dt, f0, t0, amp, 0.0, ns)
    rec_ix = np.arange(2, nx - 2, 4, dtype=np.int32)
    return dict(rid=rid, model_seed=1000 + rid, nx=nx, nz=nz, dx=float(dx), dt=float(dt), nt=nt,
                sh=int(sh), psv=int(not sh), adjoint=adjoint, sfe=sfe, src_ix=src_ix,
                src_iz=np.full(ns, 2, np.int32), stf_x=sx, stf_y=sy, stf_z=sz, rec_ix=rec_ix,
                rec_iz=np.full(len(rec_ix), 2, np.int32))


def request_cell_updates(req):
    """Cell-updates of a request: shots * nx * nz * nt * (2 with the adjoint)."""
    return len(req["src_ix"]) * req["nx"] * req["nz"] * req["nt"] * (2 if req["adjoint"] else 1)


# ---------------------------------------------------------------------------
# the benchmark workloads (one definition for the bench and the tests)
# ---------------------------------------------------------------------------
WORKLOADS = ("c0", "c1", "c2", "c3", "c4")


def workload(name, nt=0, sigma=5):
    """(problem, shots per launch, snapshot format, checkpoint interval) of a
    benchmark workload; nt > 0 shortens the run.

    c0  P-SV checkerboard 200x200 at 10 m, 4 explosive sources at 20 cells
        depth, 40 receivers at 10 cells, dt 1 ms, nt 2000, waveform misfit,
        sigma 5, FP32 snapshots; all 4 shots per launch.
    c1  P-SV on the Marmousi-shaped model, 737x240 at 12.5 m, 16 shots at
        f0 5 Hz, 368 surface receivers, dt 1 ms, nt 4000, FP32 snapshots; all
        16 shots per launch.
    c2  SH on the same generator at 1474x480 and 6.25 m (seed 2), 64 shots at
        f0 8 Hz and 737 receivers on the seafloor (row 14: SH waves do not
        enter the water layer, where mu = 0), dt 0.5 ms, nt 8000, 16-bit (I16)
        snapshots every 10 steps, all 64 shots' snapshots resident (145 GB).
    c4  P-SV on the generator at 4096x2048 and 5 m (seed 4), 256 shots at
        f0 12 Hz, 2048 receivers, dt 0.4 ms, nt 10000, envelope misfit, 16-bit
        (I16) snapshots; four shots per launch keep all of their snapshots
        resident (268 GB), so the adjoint recomputes nothing.
    The 16-bit format is I16 (fixed point per row segment), whose gradients
    stay within 1e-4 of FP64; FP16 and BF16 cost the same and are reported
    by the accuracy checks.
    (c3, the mixed-size request load, is built by `serving_request`.)"""
    if name == "c0":
        P = checker(nt=nt or 2000, sigma=sigma, explosive=True)
        return P, 4, invhip.INV_SNAP_F32, 0
    if name == "c1":
        return marmousi(nt=nt or 4000, sigma=sigma), 16, invhip.INV_SNAP_F32, 0
    if name == "c2":
        P = marmousi(nx=1474, nz=480, dx=6.25, dt=5e-4, nt=nt or 8000, nsrc=64, nrec=737,
                     f0=8.0, t0=0.15, seed=2, sh=1, psv=0, sigma=sigma, src_depth=14,
                     rec_depth=14, name="marmousi-sh")
        return P, 64, invhip.INV_SNAP_I16, 0
    if name == "c4":
        P = marmousi(nx=4096, nz=2048, dx=5.0, dt=4e-4, nt=nt or 10000, nsrc=256, nrec=2048,
                     f0=12.0, t0=0.1, seed=4, sigma=sigma, name="marmousi-large")
        return P.with_(misfit=1), 4, invhip.INV_SNAP_I16, 0
    raise ValueError(f"unknown workload {name!r}")


def shots(P, pick):
    """The problem restricted to the shots `pick` (indices into its sources)."""
    pick = np.asarray(pick)
    return P.with_(src_ix=P.src_ix[pick], src_iz=P.src_iz[pick], stf_x=P.stf_x[pick],
                   stf_y=P.stf_y[pick], stf_z=P.stf_z[pick],
                   src_type=None if P.src_type is None else P.src_type[pick])


# ---------------------------------------------------------------------------
# the reference's file inputs: model points, config directories
# ---------------------------------------------------------------------------
def gauss_weights(n, sigma):
    """The one-sided Gaussian weights of filter/gaussian.h:4-8 in FP32:
    g(m) = exp(-m^2 / (2 sigma^2)) / (sqrt(2 pi) sigma) for m = 0 .. n-1 (the
    reference's int - size_t argument wraps for negative m, whose weight is
    then 0)."""
    m = np.arange(n, dtype=np.float32)
    s = np.float32(sigma)
    pi = np.float32(3.1415927)
    c = np.float32(1) / (np.sqrt(np.float32(2) * pi) * s)
    return (c * np.exp(-m * m / (np.float32(2) * s * s))).astype(np.float32)


def one_sided_gauss(field, nx, nz, sigma):
    """GaussianFilter::apply (filter/gaussian.h:9-56) on a [nx][nz] field:
    out(i, j) = sum over n <= i, m <= j of g(i-n) g(j-m) f(n, m), divided by
    the product of the one-sided weight sums; FP64 sums of FP32 weights."""
    f = np.asarray(field, np.float64).reshape(nx, nz)
    w = gauss_weights(max(nx, nz), sigma).astype(np.float64)
    cut = int(np.nonzero(w)[0].max()) + 1
    tmp = np.zeros_like(f)
    for m in range(min(cut, nx)):
        tmp[m:, :] += w[m] * f[:nx - m, :]
    out = np.zeros_like(f)
    for m in range(min(cut, nz)):
        out[:, m:] += w[m] * tmp[:, :nz - m]
    sx = np.cumsum(w[:nx].astype(np.float32)).astype(np.float64)
    sz = np.cumsum(w[:nz].astype(np.float32)).astype(np.float64)
    return (out / (sx[:, None] * sz[None, :])).astype(np.float32).ravel()


def checker_gapped(cx, cz, nx, nz, dx, dz, vp0, vs0, rho0, dvp, dvs, drho, sigma):
    """The checkerboard of util/generate_checker.cu:4-39 (its model points and
    optional smoothing, generate_checker.cu:44-70): squares of width
    wx = dx nx 2/3 / (cx+1) (and wz likewise) separated by gaps of wx/2,
    the first starting at wx, alternating +-(dvp, dvs, drho) around the
    background; fields with a perturbation above 0.1 are smoothed with the
    one-sided Gaussian of sigma cells.  Returns the point arrays x, z, vp,
    vs, rho in the order of the reference's example files (x fastest)."""
    f32 = np.float32
    i = np.arange(nx)[:, None]
    j = np.arange(nz)[None, :]
    x = (i * f32(dx)).astype(f32) + np.zeros((1, nz), f32)
    z = (j * f32(dz)).astype(f32) + np.zeros((nx, 1), f32)
    # float arithmetic in the kernel's order; the division by 1.5 (a double
    # literal) and the size_t conversion (0 for the negative values left of
    # the first square)
    wx = f32(f32(f32(f32(dx) * f32(nx)) * f32(2)) / f32(3)) / f32(cx + 1)
    wz = f32(f32(f32(f32(dz) * f32(nz)) * f32(2)) / f32(3)) / f32(cz + 1)
    tx = ((x - wx) / wx).astype(f32).astype(np.float64) / 1.5
    tz = ((z - wz) / wz).astype(f32).astype(np.float64) / 1.5
    idx = np.where(tx > 0, np.floor(tx), 0).astype(np.int64)
    idz = np.where(tz > 0, np.floor(tz), 0).astype(np.int64)
    rx = (x - wx).astype(f32).astype(np.float64) - (idx * wx).astype(f32).astype(np.float64) * 1.5
    rz = (z - wz).astype(f32).astype(np.float64) - (idz * wz).astype(f32).astype(np.float64) * 1.5
    inside = (rx > 0) & (rx < wx) & (rz > 0) & (rz < wz)
    same = (idx % 2) == (idz % 2)

    def field(v0, dv):
        a = np.where(inside, np.where(same, f32(v0 + dv), f32(v0 - dv)), f32(v0)).astype(f32)
        if sigma and dv > 0.1:
            a = one_sided_gauss(a, nx, nz, sigma).reshape(nx, nz)
        return a

    out = [x, z, field(vp0, dvp), field(vs0, dvs), field(rho0, drho)]
    # x fastest: flat index j * nx + i
    return tuple(np.ascontiguousarray(a.T, f32).ravel() for a in out)


def write_bin(path, a):
    """`.bin` file: int32 value count, then float32 values (solver/solver.h:109-120)."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    with open(path, "wb") as f:
        f.write(np.int32(a.size).tobytes())
        f.write(a.tobytes())


CONFIG_DEFAULTS = dict(
    mode=0, clean=1, nthread=0, solver=0, model_type=1, trace_type=0, wfe=0, wae=0,
    filter=0, trace_file=0, optimizer=0, inv_iteration=1, inv_cycle=0, inv_sharpen=0.0,
    nlcg_type=1, nlcg_thresh=0.0, lbfgs_mem=5, lbfgs_thresh=0.0, ls_step=10, ls_step_max=0.5,
    ls_step_init=0.05, ls_thresh=1.2)


def write_config(P, path, **over):
    """A configuration directory the command-line driver (inv.out, inv.cu's
    modes) runs: config.ini (the keys of lib/config.h, from the problem plus
    CONFIG_DEFAULTS and `over`), sources.dat (x z type f0 t0 angle amp) and
    stations.dat (x z) in metres (lib/config.h:45-84), and model_true /
    model_init with proc000000_{x,z,lambda,mu,rho}.bin (model_type 1) on the
    grid nodes, z fastest.  The driver derives the same nx, nz from the point
    count and the extent (solver/fdm.h:356-389)."""
    from pathlib import Path

    root = Path(path)
    root.mkdir(parents=True, exist_ok=True)
    gx, gz, _, _ = invhip.grid_dims(P.nx * P.nz, (P.nx - 1) * float(P.dx),
                                    (P.nz - 1) * float(P.dz))
    if (gx, gz) != (P.nx, P.nz):
        raise ValueError(f"the driver would derive a {gx}x{gz} grid, not {P.nx}x{P.nz}")
    cfg = dict(CONFIG_DEFAULTS)
    cfg.update(sh=P.sh, psv=P.psv, nt=P.nt, dt=float(P.dt), sfe=P.sfe,
               abs_bottom=P.abs_bottom, abs_right=P.abs_right, abs_top=P.abs_top,
               abs_left=P.abs_left, abs_width=P.abs_width, abs_alpha=P.abs_alpha,
               misfit=P.misfit, filter=0 if P.sigma > 0 else 1, filter_sigma=P.sigma,
               inv_lambda=P.inv_lambda, inv_mu=P.inv_mu, inv_rho=P.inv_rho)
    cfg.update(over)
    lines = ["[global]"] + [f"{k} = {float(v)!r}" if isinstance(v, (float, np.floating))
                            else f"{k} = {int(v)}" for k, v in cfg.items()]
    (root / "config.ini").write_text("\n".join(lines) + "\n")
    meta = P.meta
    types = P.src_type if P.src_type is not None else np.zeros(P.nsrc, np.int32)
    rows = ["# x | z | type | f0 | t0 | angle | amp"]
    num = lambda v: repr(float(v))
    for s in range(P.nsrc):
        rows.append(f"{num(P.src_ix[s] * float(P.dx))} {num(P.src_iz[s] * float(P.dz))} "
                    f"{int(types[s])} {num(meta['f0'])} {num(meta['t0'])} 0.0 {num(meta['amp'])}")
    (root / "sources.dat").write_text("\n".join(rows) + "\n")
    rows = ["# x | z"] + [f"{num(P.rec_ix[r] * float(P.dx))} {num(P.rec_iz[r] * float(P.dz))}"
                          for r in range(P.nrec)]
    (root / "stations.dat").write_text("\n".join(rows) + "\n")
    i = np.arange(P.nx)[:, None]
    j = np.arange(P.nz)[None, :]
    x = (i * np.float32(P.dx) + np.zeros((1, P.nz), np.float32)).astype(np.float32).ravel()
    z = (j * np.float32(P.dz) + np.zeros((P.nx, 1), np.float32)).astype(np.float32).ravel()
    for name, (lam, mu, rho) in (("model_true", (P.lam, P.mu, P.rho)),
                                 ("model_init", (P.lam0, P.mu0, P.rho0))):
        d = root / name
        d.mkdir(exist_ok=True)
        for c, a in (("x", x), ("z", z), ("lambda", lam), ("mu", mu), ("rho", rho)):
            write_bin(d / f"proc000000_{c}.bin", a)
    return root

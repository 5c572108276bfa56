"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.

Also holds the reference's file readers/writers restated in numpy (the `.bin`
model layout of solver/solver.h:109-120 and the SU trace layout of
solver/solver.h:38-108), used to load the reference's golden fixtures.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Sim(C.Structure):
    """orc_sim of oracle.h."""

    _fields_ = [
        ("nx", C.c_int), ("nz", C.c_int),
        ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
        ("nt", C.c_int), ("sh", C.c_int), ("psv", C.c_int), ("sfe", C.c_int),
        ("adj", C.c_int), ("trace_type", C.c_int),
        ("inv_lambda", C.c_int), ("inv_mu", C.c_int), ("inv_rho", C.c_int),
        ("lam", C.c_void_p), ("mu", C.c_void_p), ("rho", C.c_void_p), ("ab", C.c_void_p),
        ("nsrc", C.c_int), ("src_ix", C.c_void_p), ("src_iz", C.c_void_p),
        ("stf_x", C.c_void_p), ("stf_y", C.c_void_p), ("stf_z", C.c_void_p),
        ("nrec", C.c_int), ("rec_ix", C.c_void_p), ("rec_iz", C.c_void_p),
        ("src_type", C.c_void_p),
    ]


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P = C.POINTER
        L.orc_grid_dims.argtypes = [C.c_int, C.c_float, C.c_float, P(C.c_int), P(C.c_int),
                                    P(C.c_float), P(C.c_float)]
        L.orc_absbound.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                   C.c_int, C.c_int, _f32p]
        L.orc_calc_idx.argtypes = [C.c_float, C.c_float, C.c_int]
        L.orc_calc_idx.restype = C.c_int
        L.orc_ricker.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, _f32p]
        L.orc_source.argtypes = [C.c_int, C.c_float, _f32p, _f32p, _f32p, _f32p]
        L.orc_mesh2grid.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int,
                                    _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_vps2lm.argtypes = [C.c_int, _f32p, _f32p, _f32p]
        L.orc_lm2vps.argtypes = [C.c_int, _f32p, _f32p, _f32p]
        for name in ("orc_forward_f32", "orc_forward_f64", "orc_adjoint_f32", "orc_adjoint_f64",
                     "orc_gradient_f32", "orc_gradient_f64"):
            getattr(L, name).restype = C.c_int if "gradient" not in name else C.c_double
        L.orc_forward_f32.argtypes = [P(Sim), C.c_int, C.c_int] + [C.c_void_p] * 5
        L.orc_forward_f64.argtypes = [P(Sim), C.c_int, C.c_int] + [C.c_void_p] * 5
        L.orc_adjoint_f32.argtypes = [P(Sim)] + [C.c_void_p] * 8
        L.orc_adjoint_f64.argtypes = [P(Sim)] + [C.c_void_p] * 8
        L.orc_gradient_f32.argtypes = [P(Sim), C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 3
        L.orc_gradient_f64.argtypes = [P(Sim), C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 3
        L.orc_waveform_f32.argtypes = [C.c_int, C.c_float, _f32p, _f32p, _f32p]
        L.orc_waveform_f32.restype = C.c_double
        L.orc_waveform_f64.argtypes = [C.c_int, C.c_double, _f64p, _f64p, _f64p]
        L.orc_waveform_f64.restype = C.c_double
        L.orc_envelope_f32.argtypes = [C.c_int, C.c_float, _f32p, _f32p, _f32p]
        L.orc_envelope_f32.restype = C.c_double
        L.orc_envelope_f64.argtypes = [C.c_int, C.c_double, _f64p, _f64p, _f64p]
        L.orc_envelope_f64.restype = C.c_double
        L.orc_hilbert_f64.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.orc_gauss_f32.argtypes = [C.c_int, C.c_int, C.c_int, _f32p]
        L.orc_gauss_f64.argtypes = [C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_snap_fields.argtypes = [P(Sim)]
        L.orc_state_fields.argtypes = [P(Sim)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# setup helpers (float, as the reference computes them)
# ---------------------------------------------------------------------------
def grid_dims(npt, xmax, zmax):
    nx, nz, dx, dz = C.c_int(), C.c_int(), C.c_float(), C.c_float()
    rc = lib().orc_grid_dims(npt, xmax, zmax, C.byref(nx), C.byref(nz), C.byref(dx), C.byref(dz))
    if rc:
        raise ValueError("invalid model geometry")
    return nx.value, nz.value, np.float32(dx.value), np.float32(dz.value)


def absbound(nx, nz, width, alpha, left=1, right=1, bottom=1, top=1):
    ab = np.empty(nx * nz, np.float32)
    lib().orc_absbound(nx, nz, width, alpha, left, right, bottom, top, ab)
    return ab


def calc_idx(coord, L, n):
    return lib().orc_calc_idx(float(coord), float(L), int(n))


def ricker(nt, dt, f0, t0):
    stf = np.empty(nt, np.float32)
    lib().orc_ricker(nt, dt, f0, t0, stf)
    return stf


def source(nt, dt, f0, t0, angle, amp):
    cfg = np.array([f0, t0, angle, amp], np.float32)
    x, y, z = (np.empty(nt, np.float32) for _ in range(3))
    lib().orc_source(nt, dt, cfg, x, y, z)
    return x, y, z


def vps2lm(vp, vs, rho):
    lam, mu = vp.astype(np.float32).copy(), vs.astype(np.float32).copy()
    lib().orc_vps2lm(lam.size, lam, mu, np.ascontiguousarray(rho, np.float32))
    return lam, mu


def mesh2grid(nx, nz, dx, dz, xmax, zmax, px, pz, pvp, pvs, prho):
    f = lambda a: np.ascontiguousarray(a, np.float32)
    lam, mu, rho = (np.zeros(nx * nz, np.float32) for _ in range(3))
    dmax = np.float32(xmax) * np.float32(xmax) + np.float32(zmax) * np.float32(zmax)
    lib().orc_mesh2grid(nx, nz, dx, dz, float(dmax), len(px), f(px), f(pz), f(pvp), f(pvs),
                        f(prho), lam, mu, rho)
    return lam, mu, rho


# ---------------------------------------------------------------------------
# simulation wrapper
# ---------------------------------------------------------------------------
class Problem:
    """A simulation setup for the oracle (keeps every array alive)."""

    def __init__(self, nx, nz, dx, dz, dt, nt, lam, mu, rho, ab, src_ix, src_iz, stf_x,
                 stf_y, stf_z, rec_ix, rec_iz, sh=1, psv=0, sfe=10, adj=1, trace_type=0,
                 inv_lambda=1, inv_mu=1, inv_rho=1, src_type=None):
        f = lambda a: np.ascontiguousarray(a, np.float32).ravel()
        i = lambda a: np.ascontiguousarray(a, np.int32).ravel()
        self.nx, self.nz, self.nt = int(nx), int(nz), int(nt)
        self.dx, self.dz, self.dt = np.float32(dx), np.float32(dz), np.float32(dt)
        self.lam, self.mu, self.rho, self.ab = f(lam), f(mu), f(rho), f(ab)
        self.src_ix, self.src_iz, self.rec_ix, self.rec_iz = i(src_ix), i(src_iz), i(rec_ix), i(rec_iz)
        self.nsrc, self.nrec = len(self.src_ix), len(self.rec_ix)
        self.stf_x = f(stf_x).reshape(-1)
        self.stf_y = f(stf_y).reshape(-1)
        self.stf_z = f(stf_z).reshape(-1)
        self.sh, self.psv, self.sfe, self.adj, self.trace_type = sh, psv, sfe, adj, trace_type
        self.inv = (inv_lambda, inv_mu, inv_rho)
        # source mechanism per source: 0 point force, 1 explosive (see oracle.h)
        self.src_type = None if src_type is None else i(src_type)
        self.sim = self._make()

    def _make(self):
        s = Sim()
        s.nx, s.nz, s.dx, s.dz, s.dt, s.nt = self.nx, self.nz, self.dx, self.dz, self.dt, self.nt
        s.sh, s.psv, s.sfe, s.adj, s.trace_type = self.sh, self.psv, self.sfe, self.adj, self.trace_type
        s.inv_lambda, s.inv_mu, s.inv_rho = self.inv
        s.lam, s.mu, s.rho, s.ab = (_ptr(a) for a in (self.lam, self.mu, self.rho, self.ab))
        s.nsrc, s.src_ix, s.src_iz = self.nsrc, _ptr(self.src_ix), _ptr(self.src_iz)
        s.stf_x, s.stf_y, s.stf_z = _ptr(self.stf_x), _ptr(self.stf_y), _ptr(self.stf_z)
        s.nrec, s.rec_ix, s.rec_iz = self.nrec, _ptr(self.rec_ix), _ptr(self.rec_iz)
        s.src_type = None if self.src_type is None else _ptr(self.src_type)
        return s

    @property
    def n(self):
        return self.nx * self.nz

    @property
    def ncomp(self):
        return (1 if self.sh else 0) + (2 if self.psv else 0)

    @property
    def nsnap(self):
        return (2 if self.sh else 0) + (4 if self.psv else 0)

    @property
    def nstate(self):
        return (4 if self.sh else 0) + (7 if self.psv else 0)

    def forward(self, isrc, nsteps=-1, dtype=np.float32, snapshots=False):
        """Returns (traces {comp: [nrec][nt]}, snapshots or None, state [nstate][N])."""
        L = lib()
        tr = {c: np.zeros((self.nrec, self.nt), dtype) for c in "xyz"}
        nsfe = self.nt // self.sfe if self.sfe else 0
        snap = np.zeros((nsfe, self.nsnap, self.n), dtype) if snapshots else None
        state = np.zeros((self.nstate, self.n), dtype)
        fn = L.orc_forward_f32 if dtype == np.float32 else L.orc_forward_f64
        rc = fn(C.byref(self.sim), int(isrc), int(nsteps), _ptr(tr["x"]), _ptr(tr["y"]),
                _ptr(tr["z"]), _ptr(snap), _ptr(state))
        if rc:
            raise RuntimeError("oracle forward failed")
        return tr, snap, state

    def adjoint(self, adstf, snap, kernels=None, dtype=np.float32):
        """adstf {comp: [nrec][nt]}; returns (k_lambda, k_mu, k_rho, state)."""
        L = lib()
        z = np.zeros((self.nrec, self.nt), dtype)
        a = {c: np.ascontiguousarray(adstf.get(c, z), dtype) for c in "xyz"}
        if kernels is None:
            kernels = [np.zeros(self.n, dtype) for _ in range(3)]
        state = np.zeros((self.nstate, self.n), dtype)
        fn = L.orc_adjoint_f32 if dtype == np.float32 else L.orc_adjoint_f64
        rc = fn(C.byref(self.sim), _ptr(a["x"]), _ptr(a["y"]), _ptr(a["z"]),
                _ptr(np.ascontiguousarray(snap, dtype)), *(_ptr(k) for k in kernels), _ptr(state))
        if rc:
            raise RuntimeError("oracle adjoint failed")
        return kernels[0], kernels[1], kernels[2], state

    def gradient(self, obs, misfit_type=0, sigma=0, kernel=True, dtype=np.float32):
        """obs [nsrc][ncomp][nrec][nt]; returns (misfit, k_lambda, k_mu, k_rho)."""
        L = lib()
        obs = np.ascontiguousarray(obs, dtype)
        ks = [np.zeros(self.n, dtype) for _ in range(3)]
        fn = L.orc_gradient_f32 if dtype == np.float32 else L.orc_gradient_f64
        f = fn(C.byref(self.sim), _ptr(obs), misfit_type, sigma, 1 if kernel else 0,
               *(_ptr(k) for k in ks))
        return f, ks[0], ks[1], ks[2]


def waveform(syn, obs, dt, dtype=np.float32):
    syn = np.ascontiguousarray(syn, dtype)
    obs = np.ascontiguousarray(obs, dtype)
    ad = np.zeros_like(syn)
    fn = lib().orc_waveform_f32 if dtype == np.float32 else lib().orc_waveform_f64
    chi = fn(len(syn), dt, syn, obs, ad)
    return chi, ad


def envelope(syn, obs, dt, dtype=np.float32):
    syn = np.ascontiguousarray(syn, dtype)
    obs = np.ascontiguousarray(obs, dtype)
    ad = np.zeros_like(syn)
    fn = lib().orc_envelope_f32 if dtype == np.float32 else lib().orc_envelope_f64
    chi = fn(len(syn), dt, syn, obs, ad)
    return chi, ad


def hilbert(x):
    x = np.ascontiguousarray(x, np.float64)
    re, im = np.zeros_like(x), np.zeros_like(x)
    lib().orc_hilbert_f64(len(x), x, re, im)
    return re + 1j * im


def gauss(field, nx, nz, sigma, dtype=np.float32):
    a = np.ascontiguousarray(field, dtype).copy().ravel()
    fn = lib().orc_gauss_f32 if dtype == np.float32 else lib().orc_gauss_f64
    fn(nx, nz, sigma, a)
    return a


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


# ---------------------------------------------------------------------------
# reference file formats (numpy restatement)
# ---------------------------------------------------------------------------
def read_bin(path):
    """`.bin`: int32 npt, then npt float32 (solver/solver.h:109-120)."""
    raw = Path(path).read_bytes()
    npt = int(np.frombuffer(raw[:4], np.int32)[0])
    return np.frombuffer(raw[4:4 + 4 * npt], np.float32).copy()


def read_su(path, nt):
    """`.su`: per receiver a 240-byte header then nt float32 (solver/solver.h:38-108)."""
    raw = np.fromfile(path, np.uint8)
    rec = 240 + 4 * nt
    nrec = raw.size // rec
    a = raw[: nrec * rec].reshape(nrec, rec)
    hdr = a[:, :240].copy()
    data = a[:, 240:].copy().view(np.float32).reshape(nrec, nt)
    return hdr, data


def read_dat(path, ncol):
    """sources.dat / stations.dat: whitespace floats, non-numeric lines skipped
    (lib/config.h:45-84)."""
    rows = []
    for line in Path(path).read_text().split("\n"):
        line = line.lstrip(" \t")
        if line != line.rstrip():
            continue  # the stream must end exactly at the last value
        toks = line.split()
        if len(toks) < ncol:
            continue
        try:
            vals = [float(t) for t in toks[:ncol]]
        except ValueError:
            continue
        if len(toks) != ncol:
            continue
        rows.append(vals)
    return np.array(rows, np.float32).reshape(-1, ncol)

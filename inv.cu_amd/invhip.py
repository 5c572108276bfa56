"""ctypes binding of libinv_hip.so (include/inv_hip.h).

Every call goes through the C ABI; device memory is owned by the library's
context.  If the shared library is missing this module raises at load time:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libinv_hip.so"

INV_OK = 0
INV_LAMBDA, INV_MU, INV_RHO, INV_K_LAMBDA, INV_K_MU, INV_K_RHO, INV_OUT, INV_ADSTF, INV_OBS = range(9)
INV_SNAP_F32, INV_SNAP_F16, INV_SNAP_BF16, INV_SNAP_I16 = 0, 1, 2, 3
INV_FWD_SNAPSHOTS = 1
INV_PROF_VELOCITY, INV_PROF_STRESS, INV_PROF_INTERACT, INV_PROF_STEP = 0, 1, 2, 3
INV_PROF_FWD_RUN, INV_PROF_ADJ_RUN, INV_PROF_PAIR, INV_PROF_RESIDENT = 4, 5, 6, 7
INV_PROF_TRIPLE = 8
INV_VARIANT_FUSED, INV_VARIANT_SPLIT, INV_VARIANT_FUSED_LINEAR = 0, 1, 2
INV_MATH_EXACT, INV_MATH_FAST = 0, 1

STATUS = {0: "INV_OK", -1: "INV_ERR_ARG", -2: "INV_ERR_HIP", -3: "INV_ERR_STATE",
          -4: "INV_ERR_NOMEM", -5: "INV_ERR_UNSUPPORTED"}


class GridDesc(C.Structure):
    """inv_grid_desc of include/inv_hip.h."""

    _fields_ = [
        ("nx", C.c_int), ("nz", C.c_int),
        ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
        ("nt", C.c_int), ("sh", C.c_int), ("psv", C.c_int), ("adj", C.c_int),
        ("trace_type", C.c_int), ("sfe", C.c_int), ("abs_width", C.c_int),
        ("abs_alpha", C.c_float), ("abs_left", C.c_int), ("abs_right", C.c_int),
        ("abs_top", C.c_int), ("abs_bottom", C.c_int),
        ("inv_lambda", C.c_int), ("inv_mu", C.c_int), ("inv_rho", C.c_int),
        ("batch", C.c_int),
    ]


# Every exported function and its argument types (checked against the header
# by tests/test_capi.py).
_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_sz = C.c_size_t
SIGNATURES = {
    "inv_grid_dims": [_i, _f, _f, _vp, _vp, _vp, _vp],
    "inv_calc_idx": [_f, _f, _i],
    "inv_source_ricker": [_i, _f, _f, _f, _f, _f, _vp, _vp, _vp],
    "inv_device_count": [_vp],
    "inv_ctx_create": [_i, _vp],
    "inv_ctx_destroy": [_vp],
    "inv_last_error": [_vp],
    "inv_sync": [_vp],
    "inv_mem_info": [_vp, _vp, _vp],
    "inv_grid_set": [_vp, _vp],
    "inv_model_upload_lmr": [_vp, _vp, _vp, _vp],
    "inv_model_download_lmr": [_vp, _vp, _vp, _vp],
    "inv_model_from_points": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _f, _f, _i],
    "inv_model_to_vps": [_vp, _vp, _vp, _vp],
    "inv_sources_set": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
    "inv_receivers_set": [_vp, _i, _vp, _vp],
    "inv_sources_type": [_vp, _vp],
    "inv_forward": [_vp, _i, _i, _i],
    "inv_adjoint": [_vp, _i, _i],
    "inv_kernels_zero": [_vp],
    "inv_kernels_get": [_vp, _vp, _vp, _vp],
    "inv_kernels_set": [_vp, _vp, _vp, _vp],
    "inv_traces_get": [_vp, _i, _vp],
    "inv_adstf_set": [_vp, _i, _vp],
    "inv_adstf_get": [_vp, _i, _vp],
    "inv_obs_set": [_vp, _i, _vp],
    "inv_obs_get": [_vp, _i, _vp],
    "inv_obs_from_traces": [_vp, _i, _i],
    "inv_misfit": [_vp, _i, _i, _i, _vp, _vp],
    "inv_gradient": [_vp, _i, _i, _i, _vp],
    "inv_filter_gauss1s": [_vp, _vp, _i, _i, _i],
    "inv_vec_axpby": [_vp, _vp, _f, _vp, _f, _vp, _sz],
    "inv_vec_scale": [_vp, _vp, _f, _vp, _sz],
    "inv_vec_mul": [_vp, _vp, _vp, _vp, _sz],
    "inv_vec_copy": [_vp, _vp, _vp, _sz],
    "inv_vec_fill": [_vp, _vp, _f, _sz],
    "inv_vec_dot_f64": [_vp, _vp, _vp, _sz, _vp],
    "inv_vec_amax": [_vp, _vp, _sz, _vp],
    "inv_field_ptr": [_vp, _i, _vp],
    "inv_alloc": [_vp, _sz, _vp],
    "inv_free": [_vp, _vp],
    "inv_memcpy_h2d": [_vp, _vp, _vp, _sz],
    "inv_memcpy_d2h": [_vp, _vp, _vp, _sz],
    "inv_state_get": [_vp, _i, _vp],
    "inv_state_set": [_vp, _i, _vp],
    "inv_step_velocity": [_vp, _i, _i, _i],
    "inv_step_stress": [_vp, _i],
    "inv_snapshots_get": [_vp, _i, _vp],
    "inv_set_step_limit": [_vp, _i],
    "inv_set_snapshots": [_vp, _i, _i],
    "inv_prof_enable": [_vp, _i],
    "inv_prof_reset": [_vp],
    "inv_prof_read": [_vp, _i, _vp, _vp],
    "inv_set_variant": [_vp, _i, _i],
    "inv_set_tile": [_vp, _i, _i],
    "inv_set_pair": [_vp, _i, _i],
    "inv_get_pair": [_vp, _vp, _vp],
    "inv_set_triple": [_vp, _i],
    "inv_get_triple": [_vp, _vp],
    "inv_set_walk": [_vp, _i],
    "inv_get_walk": [_vp, _vp],
    "inv_set_persistent": [_vp, _i],
    "inv_get_persistent": [_vp, _vp],
    "inv_autotune": [_vp, _i, _i, _i, _vp, _vp, _vp],
    "inv_forward_begin": [_vp, _i, _i, _i],
    "inv_adjoint_begin": [_vp, _i, _i],
    "inv_run_advance": [_vp, _i, _vp],
    "inv_misfit_trace": [_vp, _i, _vp, _vp, _vp, _vp],
    "inv_misfit_traces_noalloc": [_vp, _vp, _vp, _vp, _i, _vp, _vp],
    "inv_set_math": [_vp, _i],
    "inv_set_graphs": [_vp, _i],
    "inv_server_create": [_vp, _vp],
    "inv_server_destroy": [_vp],
    "inv_server_last_error": [_vp],
    "inv_server_model_add": [_vp, _i, _i, _vp, _vp, _vp, _vp],
    "inv_server_model_remove": [_vp, _i],
    "inv_server_submit": [_vp, _vp, _vp],
    "inv_server_wait": [_vp, C.c_longlong, C.c_double, _vp],
    "inv_server_poll": [_vp, C.c_longlong, _vp],
    "inv_server_stats_get": [_vp, _vp],
}
_CHAR_RESULT = {"inv_last_error", "inv_server_last_error"}

_lib = None


def load():
    """Load libinv_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                               "(make -C inv.cu_amd)")
        L = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name in _CHAR_RESULT else C.c_int
        _lib = L
    return _lib


def grid_dims(npt, xmax, zmax):
    nx, nz, dx, dz = C.c_int(), C.c_int(), C.c_float(), C.c_float()
    rc = load().inv_grid_dims(npt, xmax, zmax, C.byref(nx), C.byref(nz), C.byref(dx), C.byref(dz))
    if rc != INV_OK:
        raise InvError(rc, "invalid model geometry")
    return nx.value, nz.value, np.float32(dx.value), np.float32(dz.value)


def calc_idx(coord, length, n):
    return load().inv_calc_idx(float(coord), float(length), int(n))


def source_ricker(nt, dt, f0, t0=0.0, angle=0.0, amp=1.0):
    """Wavelets (stf_x, stf_y, stf_z) of one source, each float32 [nt]."""
    out = [np.empty(nt, np.float32) for _ in range(3)]
    rc = load().inv_source_ricker(nt, dt, f0, t0, angle, amp, *(_p(a) for a in out))
    if rc != INV_OK:
        raise InvError(rc, "bad source parameters")
    return out


def device_count() -> int:
    n = C.c_int(0)
    load().inv_device_count(C.byref(n))
    return n.value


class InvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context of the library (inv_ctx)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.inv_ctx_create(device, C.byref(h))
        if rc != INV_OK:
            raise InvError(rc, f"inv_ctx_create({device}) failed")
        self.h = h
        self.desc = None

    def close(self):
        if getattr(self, "h", None):
            self.L.inv_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc):
        if rc != INV_OK:
            raise InvError(rc, self.L.inv_last_error(self.h).decode())

    # ---- setup ----
    def grid_set(self, **kw):
        d = GridDesc()
        defaults = dict(sh=0, psv=0, adj=1, trace_type=0, sfe=10, abs_width=20, abs_alpha=0.015,
                        abs_left=1, abs_right=1, abs_top=1, abs_bottom=1, inv_lambda=1, inv_mu=1,
                        inv_rho=1, batch=1)
        defaults.update(kw)
        for k, v in defaults.items():
            setattr(d, k, v)
        self._chk(self.L.inv_grid_set(self.h, C.byref(d)))
        self.desc = d
        self._ckpt = 0  # inv_grid_set resets snapshot storage to FP32, no checkpoints
        return d

    @property
    def N(self):
        return self.desc.nx * self.desc.nz

    @property
    def ncomp(self):
        return (1 if self.desc.sh else 0) + (2 if self.desc.psv else 0)

    @property
    def nsnap(self):
        return (2 if self.desc.sh else 0) + (4 if self.desc.psv else 0)

    @property
    def nstate(self):
        return (4 if self.desc.sh else 0) + (7 if self.desc.psv else 0)

    def model_upload(self, lam, mu, rho):
        f = lambda a: np.ascontiguousarray(a, np.float32).ravel()
        lam, mu, rho = f(lam), f(mu), f(rho)
        self._chk(self.L.inv_model_upload_lmr(self.h, _p(lam), _p(mu), _p(rho)))

    def model_download(self):
        out = [np.empty(self.N, np.float32) for _ in range(3)]
        self._chk(self.L.inv_model_download_lmr(self.h, *(_p(a) for a in out)))
        return out

    def model_from_points(self, px, pz, p1, p2, prho, xmax, zmax, model_type=0):
        """mesh2grid (+ vps2lm for model_type 0) of npt model points
        (inv_model_from_points)."""
        f = lambda a: np.ascontiguousarray(a, np.float32).ravel()
        px, pz, p1, p2, prho = f(px), f(pz), f(p1), f(p2), f(prho)
        self._chk(self.L.inv_model_from_points(self.h, len(px), _p(px), _p(pz), _p(p1), _p(p2),
                                               _p(prho), float(xmax), float(zmax),
                                               int(model_type)))

    def sources_set(self, ix, iz, stf_x, stf_y, stf_z):
        ix = np.ascontiguousarray(ix, np.int32)
        iz = np.ascontiguousarray(iz, np.int32)
        f = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1)
        sx, sy, sz = f(stf_x), f(stf_y), f(stf_z)
        self.nsrc = len(ix)
        self._chk(self.L.inv_sources_set(self.h, len(ix), _p(ix), _p(iz), _p(sx), _p(sy), _p(sz)))

    def sources_type(self, types):
        """Source mechanism per source: 0 point force, 1 explosive."""
        t = np.ascontiguousarray(types, np.int32)
        assert len(t) == self.nsrc
        self._chk(self.L.inv_sources_type(self.h, _p(t)))

    def receivers_set(self, ix, iz):
        ix = np.ascontiguousarray(ix, np.int32)
        iz = np.ascontiguousarray(iz, np.int32)
        self.nrec = len(ix)
        self._chk(self.L.inv_receivers_set(self.h, len(ix), _p(ix), _p(iz)))

    # ---- propagation ----
    def forward(self, first, count=1, snapshots=True):
        self._chk(self.L.inv_forward(self.h, first, count, INV_FWD_SNAPSHOTS if snapshots else 0))

    def adjoint(self, first, count=1):
        self._chk(self.L.inv_adjoint(self.h, first, count))

    def forward_begin(self, first, count=1, snapshots=True):
        self._chk(self.L.inv_forward_begin(self.h, first, count,
                                           INV_FWD_SNAPSHOTS if snapshots else 0))

    def adjoint_begin(self, first, count=1):
        self._chk(self.L.inv_adjoint_begin(self.h, first, count))

    def advance(self, nsteps):
        pos = C.c_int()
        self._chk(self.L.inv_run_advance(self.h, nsteps, C.byref(pos)))
        return pos.value

    def misfit_trace(self, kind, syn_dev, obs_dev, adstf_dev):
        chi = C.c_double()
        self._chk(self.L.inv_misfit_trace(self.h, kind, syn_dev, obs_dev, adstf_dev,
                                          C.byref(chi)))
        return chi.value

    def misfit_traces_noalloc(self, syn_dev, obs_dev, adstf_dev, ntr):
        """Waveform misfit of ntr device traces with no device allocation;
        returns (per-trace misfits, total)."""
        chi = np.zeros(ntr, np.float64)
        tot = C.c_double()
        self._chk(self.L.inv_misfit_traces_noalloc(self.h, C.c_void_p(syn_dev),
                                                   C.c_void_p(obs_dev), C.c_void_p(adstf_dev),
                                                   ntr, _p(chi), C.byref(tot)))
        return chi, tot.value

    def kernels_zero(self):
        self._chk(self.L.inv_kernels_zero(self.h))

    def kernels_get(self):
        out = [np.empty(self.N, np.float32) for _ in range(3)]
        self._chk(self.L.inv_kernels_get(self.h, *(_p(a) for a in out)))
        return out

    def traces_get(self, slot=0):
        a = np.empty((self.ncomp, self.nrec, self.desc.nt), np.float32)
        self._chk(self.L.inv_traces_get(self.h, slot, _p(a)))
        return a

    def adstf_set(self, slot, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == self.ncomp * self.nrec * self.desc.nt
        self._chk(self.L.inv_adstf_set(self.h, slot, _p(a)))

    def adstf_get(self, slot=0):
        a = np.empty((self.ncomp, self.nrec, self.desc.nt), np.float32)
        self._chk(self.L.inv_adstf_get(self.h, slot, _p(a)))
        return a

    def obs_set(self, isrc, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == self.ncomp * self.nrec * self.desc.nt
        self._chk(self.L.inv_obs_set(self.h, isrc, _p(a)))

    def obs_get(self, isrc):
        a = np.empty((self.ncomp, self.nrec, self.desc.nt), np.float32)
        self._chk(self.L.inv_obs_get(self.h, isrc, _p(a)))
        return a

    def obs_from_traces(self, first, count=1):
        self._chk(self.L.inv_obs_from_traces(self.h, first, count))

    def misfit(self, first, count=1, kind=0):
        chi = np.empty((count, self.nrec, self.ncomp), np.float64)
        tot = C.c_double()
        self._chk(self.L.inv_misfit(self.h, first, count, kind, _p(chi), C.byref(tot)))
        return tot.value, chi

    def gradient(self, misfit_type=0, sigma=0, kernel=True):
        f = C.c_double()
        self._chk(self.L.inv_gradient(self.h, misfit_type, sigma, 1 if kernel else 0, C.byref(f)))
        return f.value

    def sync(self):
        self._chk(self.L.inv_sync(self.h))

    # ---- device memory / vectors ----
    def field_ptr(self, field):
        p = C.c_void_p()
        self._chk(self.L.inv_field_ptr(self.h, field, C.byref(p)))
        return p.value

    def alloc(self, nbytes):
        p = C.c_void_p()
        self._chk(self.L.inv_alloc(self.h, nbytes, C.byref(p)))
        return p.value

    def free(self, ptr):
        self._chk(self.L.inv_free(self.h, ptr))

    def h2d(self, dev, a):
        a = np.ascontiguousarray(a)
        self._chk(self.L.inv_memcpy_h2d(self.h, dev, _p(a), a.nbytes))

    def d2h(self, dev, shape, dtype=np.float32):
        a = np.empty(shape, dtype)
        self._chk(self.L.inv_memcpy_d2h(self.h, _p(a), dev, a.nbytes))
        return a

    def filter_gauss(self, dev, nx, nz, sigma):
        self._chk(self.L.inv_filter_gauss1s(self.h, dev, nx, nz, sigma))

    def vec_axpby(self, c, ka, a, kb, b, n):
        self._chk(self.L.inv_vec_axpby(self.h, c, ka, a, kb, b, n))

    def vec_scale(self, c, ka, a, n):
        self._chk(self.L.inv_vec_scale(self.h, c, ka, a, n))

    def vec_mul(self, c, a, b, n):
        self._chk(self.L.inv_vec_mul(self.h, c, a, b, n))

    def vec_copy(self, dst, src, n):
        self._chk(self.L.inv_vec_copy(self.h, dst, src, n))

    def vec_fill(self, dst, v, n):
        self._chk(self.L.inv_vec_fill(self.h, dst, v, n))

    def vec_dot(self, a, b, n):
        out = C.c_double()
        self._chk(self.L.inv_vec_dot_f64(self.h, a, b, n, C.byref(out)))
        return out.value

    def vec_amax(self, a, n):
        out = C.c_float()
        self._chk(self.L.inv_vec_amax(self.h, a, n, C.byref(out)))
        return out.value

    # ---- test hooks ----
    def state_get(self, slot=0):
        a = np.empty((self.nstate, self.N), np.float32)
        self._chk(self.L.inv_state_get(self.h, slot, _p(a)))
        return a

    def state_set(self, slot, a):
        a = np.ascontiguousarray(a, np.float32).reshape(self.nstate, self.N)
        self._chk(self.L.inv_state_set(self.h, slot, _p(a)))

    def step_velocity(self, first, count, it):
        """Velocity half-step of time step it (divS + addSTF + updateV)."""
        self._chk(self.L.inv_step_velocity(self.h, first, count, it))

    def step_stress(self, count=1):
        """Stress half-step (divV + updateS)."""
        self._chk(self.L.inv_step_stress(self.h, count))

    def snapshots_get(self, slot=0):
        ckpt = getattr(self, "_ckpt", 0)
        nslots = (ckpt if ckpt else self.desc.nt) // self.desc.sfe
        a = np.empty((nslots, self.nsnap, self.N), np.float32)
        self._chk(self.L.inv_snapshots_get(self.h, slot, _p(a)))
        return a

    def set_step_limit(self, n):
        self._chk(self.L.inv_set_step_limit(self.h, n))

    def set_snapshots(self, fmt=INV_SNAP_F32, ckpt=0):
        """Snapshot storage format (INV_SNAP_*) and checkpoint interval (0: none)."""
        self._chk(self.L.inv_set_snapshots(self.h, fmt, ckpt))
        self._ckpt = ckpt

    def set_graphs(self, on):
        self._chk(self.L.inv_set_graphs(self.h, 1 if on else 0))

    def set_math(self, mode):
        self._chk(self.L.inv_set_math(self.h, mode))

    def set_variant(self, variant, tile_x=0):
        self._chk(self.L.inv_set_variant(self.h, variant, tile_x))

    def autotune(self, first=0, count=1, steps=20):
        """Time the candidate step kernels on this shape and select the
        fastest: returns (variant, tile_rows, cols_per_lane)."""
        v, t, w = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.inv_autotune(self.h, first, count, steps, C.byref(v), C.byref(t),
                                      C.byref(w)))
        return v.value, t.value, w.value

    def set_tile(self, tile_x=8, cols_per_lane=1):
        """Fused column kernel with tile_x rows and cols_per_lane columns per lane."""
        self._chk(self.L.inv_set_tile(self.h, tile_x, cols_per_lane))

    def set_pair(self, tile_x=8, cols_per_lane=2):
        """Two steps per launch with tile_x rows and cols_per_lane columns per
        lane where the steps allow it (tile_x = 0: off)."""
        self._chk(self.L.inv_set_pair(self.h, tile_x, cols_per_lane))

    def get_pair(self):
        """(rows, columns per lane) of the two-step tile in use (rows 0: off)."""
        r, v = C.c_int(), C.c_int()
        self._chk(self.L.inv_get_pair(self.h, C.byref(r), C.byref(v)))
        return r.value, v.value

    def set_triple(self, tile_x=12):
        """Three steps per launch with tile_x (8, 12, 16 or 20) rows where the
        steps allow it (0: off)."""
        self._chk(self.L.inv_set_triple(self.h, tile_x))

    def get_triple(self):
        r = C.c_int()
        self._chk(self.L.inv_get_triple(self.h, C.byref(r)))
        return r.value

    def set_walk(self, chunk_rows=256):
        """Three steps per launch by the walking kernel, chunks of chunk_rows
        rows (a multiple of 8 in [16, 4096]; 0: off)."""
        self._chk(self.L.inv_set_walk(self.h, chunk_rows))

    def get_walk(self):
        r = C.c_int()
        self._chk(self.L.inv_get_walk(self.h, C.byref(r)))
        return r.value

    def set_persistent(self, mode=1):
        """Persistent two-step kernel (1), off (0), or on with every wait
        given up at once (2, a test hook)."""
        self._chk(self.L.inv_set_persistent(self.h, mode))

    def get_persistent(self):
        m = C.c_int()
        self._chk(self.L.inv_get_persistent(self.h, C.byref(m)))
        return m.value

    # ---- measurement ----
    def prof_enable(self, every):
        self._chk(self.L.inv_prof_enable(self.h, every))

    def prof_reset(self):
        self._chk(self.L.inv_prof_reset(self.h))

    def prof_read(self, kernel):
        n = C.c_longlong()
        ms = C.c_double()
        self._chk(self.L.inv_prof_read(self.h, kernel, C.byref(n), C.byref(ms)))
        return n.value, ms.value


# ---------------------------------------------------------------------------
# request server (inv_server_* of include/inv_hip.h)
# ---------------------------------------------------------------------------
INV_REQ_QUEUED, INV_REQ_RUNNING, INV_REQ_DONE, INV_REQ_FAILED = 0, 1, 2, 3


class ServerDesc(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_size_t), ("max_lanes", C.c_int),
                ("max_slots", C.c_int), ("max_bypass", C.c_int), ("snap_format", C.c_int),
                ("math", C.c_int), ("tile_x", C.c_int), ("plain", C.c_int)]


class Request(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
                ("dt", C.c_float), ("nt", C.c_int), ("sh", C.c_int), ("psv", C.c_int),
                ("adjoint", C.c_int), ("trace_type", C.c_int), ("sfe", C.c_int),
                ("abs_width", C.c_int), ("abs_alpha", C.c_float), ("abs_left", C.c_int),
                ("abs_right", C.c_int), ("abs_top", C.c_int), ("abs_bottom", C.c_int),
                ("model", C.c_int), ("nshots", C.c_int), ("src_ix", _vp), ("src_iz", _vp),
                ("stf_x", _vp), ("stf_y", _vp), ("stf_z", _vp), ("nrec", C.c_int),
                ("rec_ix", _vp), ("rec_iz", _vp), ("obs", _vp), ("traces", _vp),
                ("k_lambda", _vp), ("k_mu", _vp), ("k_rho", _vp)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("error", C.c_int), ("misfit", C.c_double),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("lane", C.c_int), ("batch_shots", C.c_int), ("batch_requests", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong),
                ("failed", C.c_longlong), ("batches", C.c_longlong), ("shots", C.c_longlong),
                ("cell_updates", C.c_double), ("budget", C.c_size_t),
                ("models_bytes", C.c_size_t), ("bytes_held", C.c_size_t),
                ("bytes_peak", C.c_size_t), ("bytes_actual", C.c_size_t),
                ("lanes_busy", C.c_int), ("queued", C.c_int), ("now", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Server:
    """The library's request server (inv_server): resident models, lanes with
    their own streams, continuous batching of shot-solve requests.

    submit() keeps the request's numpy arrays alive until wait() returns; the
    result arrays (traces, and kernels for adjoint requests) are allocated
    here and returned by wait()."""

    def __init__(self, device=0, hbm_budget=0, max_lanes=4, max_slots=16, max_bypass=4,
                 snap_format=INV_SNAP_F32, math=INV_MATH_FAST, tile_x=0, plain=False):
        self.L = load()
        d = ServerDesc(device, hbm_budget, max_lanes, max_slots, max_bypass, snap_format, math,
                       tile_x, 1 if plain else 0)
        h = C.c_void_p()
        rc = self.L.inv_server_create(C.byref(d), C.byref(h))
        if rc != INV_OK:
            raise InvError(rc, "inv_server_create failed")
        self.h = h
        self._live = {}

    def _chk(self, rc):
        if rc != INV_OK:
            raise InvError(rc, self.L.inv_server_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.inv_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def model_add(self, lam, mu, rho, nx, nz):
        a = [np.ascontiguousarray(x, np.float32).ravel() for x in (lam, mu, rho)]
        assert all(x.size == nx * nz for x in a)
        mid = C.c_int()
        self._chk(self.L.inv_server_model_add(self.h, nx, nz, *(_p(x) for x in a),
                                              C.byref(mid)))
        return mid.value

    def model_remove(self, mid):
        self._chk(self.L.inv_server_model_remove(self.h, mid))

    def submit(self, *, model, nx, nz, dx, dt, nt, src_ix, src_iz, stf_x, stf_y, stf_z,
               rec_ix, rec_iz, sh=0, psv=1, adjoint=False, obs=None, trace_type=0, sfe=10,
               abs_width=20, abs_alpha=0.015, abs_left=1, abs_right=1, abs_top=1,
               abs_bottom=1, dz=None, kernels=True):
        """Queue one request; returns its id."""
        ns = len(src_ix)
        ncomp = 1 if sh else 2
        keep = {
            "src_ix": np.ascontiguousarray(src_ix, np.int32),
            "src_iz": np.ascontiguousarray(src_iz, np.int32),
            "stf_x": np.ascontiguousarray(stf_x, np.float32).reshape(ns, nt),
            "stf_y": np.ascontiguousarray(stf_y, np.float32).reshape(ns, nt),
            "stf_z": np.ascontiguousarray(stf_z, np.float32).reshape(ns, nt),
            "rec_ix": np.ascontiguousarray(rec_ix, np.int32),
            "rec_iz": np.ascontiguousarray(rec_iz, np.int32),
            "traces": np.zeros((ns, ncomp, len(rec_ix), nt), np.float32),
        }
        r = Request()
        r.nx, r.nz, r.dx, r.dz, r.dt, r.nt = nx, nz, dx, dx if dz is None else dz, dt, nt
        r.sh, r.psv = (1, 0) if sh else (0, 1)
        r.adjoint = 1 if adjoint else 0
        r.trace_type, r.sfe = trace_type, sfe
        r.abs_width, r.abs_alpha = abs_width, abs_alpha
        r.abs_left, r.abs_right, r.abs_top, r.abs_bottom = abs_left, abs_right, abs_top, abs_bottom
        r.model, r.nshots, r.nrec = model, ns, len(rec_ix)
        for k in ("src_ix", "src_iz", "stf_x", "stf_y", "stf_z", "rec_ix", "rec_iz", "traces"):
            setattr(r, k, _p(keep[k]).value)
        if adjoint:
            keep["obs"] = np.ascontiguousarray(obs, np.float32).reshape(ns, ncomp, len(rec_ix), nt)
            r.obs = _p(keep["obs"]).value
            if kernels:
                for k in ("k_lambda", "k_mu", "k_rho"):
                    keep[k] = np.zeros(nx * nz, np.float32)
                    setattr(r, k, _p(keep[k]).value)
        rid = C.c_longlong()
        self._chk(self.L.inv_server_submit(self.h, C.byref(r), C.byref(rid)))
        keep["req"] = r
        self._live[rid.value] = keep
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """Wait for a request: (info dict, traces, kernels or None)."""
        info = RequestInfo()
        rc = self.L.inv_server_wait(self.h, rid, float(timeout), C.byref(info))
        if rc == INV_ERR_STATE_TIMEOUT and info.status in (INV_REQ_QUEUED, INV_REQ_RUNNING):
            raise TimeoutError(f"request {rid} still {info.status}")
        keep = self._live.pop(rid, None)
        self._chk(rc)
        ks = None
        if keep is not None and "k_lambda" in keep:
            ks = (keep["k_lambda"], keep["k_mu"], keep["k_rho"])
        return info.as_dict(), (keep["traces"] if keep else None), ks

    def poll(self, rid):
        info = RequestInfo()
        self._chk(self.L.inv_server_poll(self.h, rid, C.byref(info)))
        return info.as_dict()

    def stats(self):
        st = ServerStats()
        self._chk(self.L.inv_server_stats_get(self.h, C.byref(st)))
        return st.as_dict()


INV_ERR_STATE_TIMEOUT = -3

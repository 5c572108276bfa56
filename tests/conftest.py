"""Shared fixtures: the product package (through its C ABI) and the oracle."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import __graft_entry__ as graft  # noqa: E402

inv = graft.load_package()
import pyoracle  # noqa: E402  (the checker)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libinv_hip.so)")


def gpu_count() -> int:
    try:
        return inv.invhip.device_count()
    except Exception:
        return 0


def oracle_problem(P, model="init", adj=1, trace_type=None):
    """The oracle's view of a Problem (same arrays, same flags)."""
    lam, mu, rho = (P.lam, P.mu, P.rho) if model == "true" else (P.lam0, P.mu0, P.rho0)
    ab = pyoracle.absbound(P.nx, P.nz, P.abs_width, P.abs_alpha, P.abs_left, P.abs_right,
                           P.abs_bottom, P.abs_top)
    return pyoracle.Problem(P.nx, P.nz, P.dx, P.dz, P.dt, P.nt, lam, mu, rho, ab, P.src_ix,
                            P.src_iz, P.stf_x, P.stf_y, P.stf_z, P.rec_ix, P.rec_iz, sh=P.sh,
                            psv=P.psv, sfe=P.sfe, adj=adj,
                            trace_type=P.trace_type if trace_type is None else trace_type,
                            inv_lambda=P.inv_lambda, inv_mu=P.inv_mu, inv_rho=P.inv_rho,
                            src_type=P.src_type)


def oracle_traces(tr, P):
    """Oracle traces dict -> [ncomp][nrec][nt] in the library's component order."""
    comps = (["y"] if P.sh else []) + (["x", "z"] if P.psv else [])
    return np.stack([tr[c] for c in comps])


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="session")
def gpu():
    if gpu_count() < 1:
        pytest.skip("no GPU visible")
    return 0

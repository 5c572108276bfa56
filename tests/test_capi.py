"""libinv_hip.so loads and exports every symbol include/inv_hip.h declares;
the host-only setup helpers match the oracle bit for bit (no GPU needed)."""
import re
from pathlib import Path

import numpy as np

import pyoracle
from conftest import ROOT, inv

HEADER = ROOT / "include" / "inv_hip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(inv_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    L = inv.invhip.load()
    names = declared()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), f"{n} not exported"
        assert n in inv.invhip.SIGNATURES, f"{n} missing from the ctypes binding"
    assert set(inv.invhip.SIGNATURES) == set(names)


def test_ricker_matches_oracle_bitwise():
    for (nt, dt, f0, t0, ang, amp) in [(5000, 0.06, 0.02, 0.0, 0.0, 1e10),
                                       (4000, 1e-3, 5.0, 0.3, 0.0, 1e10),
                                       (777, 5e-4, 25.0, 0.04, 0.7, 1e9)]:
        a = inv.invhip.source_ricker(nt, dt, f0, t0, ang, amp)
        b = pyoracle.source(nt, np.float32(dt), f0, t0, ang, amp)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_calc_idx_and_grid_match_oracle():
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = int(rng.integers(5, 3000))
        L = float(rng.uniform(100, 1e5))
        c = float(rng.uniform(0, L))
        assert inv.invhip.calc_idx(c, L, n) == pyoracle.calc_idx(c, L, n)
    for npt, xm, zm in [(40401, 480000.0, 480000.0), (69611, 9200.0, 3000.0), (176880, 9200.0, 2987.5)]:
        assert inv.invhip.grid_dims(npt, xm, zm) == pyoracle.grid_dims(npt, xm, zm)


def test_context_without_device_fails_cleanly():
    if inv.invhip.device_count() > 0:
        return
    try:
        inv.invhip.Context(0)
    except inv.invhip.InvError as e:
        assert e.code == -1
    else:
        raise AssertionError("context creation should fail without a device")

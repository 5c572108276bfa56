"""The oracle against the reference's only golden output: the SH velocity
traces of example/checker (tests/golden/checker_fixture.npz, written by
tests/golden/make_checker_fixtures.py from the reference's own data files).

Setup as example/checker/config.ini: sh=1, psv=0, nt=5000, dt=0.06,
abs_width 20, abs_alpha 0.015, abs_top off, trace_type 0 (velocity),
mode 3 (adj off).  The grid comes from the model files (solver/fdm.h:356-389),
the model through mesh2grid + vps2lm, source/receiver cells through calcIdx.
"""
import io
from pathlib import Path

import numpy as np
import pytest

import pyoracle
from conftest import rel

FIX = Path(__file__).resolve().parent / "golden" / "checker_fixture.npz"


@pytest.fixture(scope="module")
def checker():
    d = np.load(FIX)
    x, z = d["model_x"], d["model_z"]
    xmax, zmax = float(x.max()), float(z.max())
    nx, nz, dx, dz = pyoracle.grid_dims(len(x), xmax, zmax)
    lam, mu, rho = pyoracle.mesh2grid(nx, nz, dx, dz, xmax, zmax, x, z, d["model_vp"],
                                      d["model_vs"], d["model_rho"])
    lam, mu = pyoracle.vps2lm(lam, mu, rho)
    tmp = Path("/tmp") / "inv_checker_dat"
    tmp.mkdir(exist_ok=True)
    (tmp / "sources.dat").write_bytes(d["sources_txt"].tobytes())
    (tmp / "stations.dat").write_bytes(d["stations_txt"].tobytes())
    src = pyoracle.read_dat(tmp / "sources.dat", 7)
    rec = pyoracle.read_dat(tmp / "stations.dat", 2)
    nt, dt = 5000, np.float32(0.06)
    ab = pyoracle.absbound(nx, nz, 20, 0.015, left=1, right=1, bottom=1, top=0)
    sx = [pyoracle.calc_idx(s[0], xmax, nx) for s in src]
    sz = [pyoracle.calc_idx(s[1], zmax, nz) for s in src]
    rx = [pyoracle.calc_idx(r[0], xmax, nx) for r in rec]
    rz = [pyoracle.calc_idx(r[1], zmax, nz) for r in rec]
    stf = [pyoracle.source(nt, dt, *s[3:7]) for s in src]
    P = pyoracle.Problem(nx, nz, dx, dz, dt, nt, lam, mu, rho, ab, sx, sz,
                         np.stack([s[0] for s in stf]), np.stack([s[1] for s in stf]),
                         np.stack([s[2] for s in stf]), rx, rz, sh=1, psv=0, sfe=10, adj=0,
                         trace_type=0)
    return P, d, (nx, nz, dx, dz, len(src), len(rec))


def test_geometry(checker):
    P, d, (nx, nz, dx, dz, nsrc, nrec) = checker
    assert (nx, nz, float(dx), float(dz)) == (201, 201, 2400.0, 2400.0)
    assert (nsrc, nrec) == (25, 132)
    assert d["vy_0"].shape[1] == 5000


@pytest.mark.parametrize("shot", [0, 12, 24])
def test_sh_traces_match_reference_golden(checker, shot):
    """FP32 and FP64 oracle reproduce vy_0000NN.su within the FP32 noise floor."""
    P, d, _ = checker
    step = int(d["rec_step"])
    gold = d[f"vy_{shot}"]
    tr, _, _ = P.forward(shot)
    got = tr["y"][::step]
    e32 = rel(got, gold)
    tr64, _, _ = P.forward(shot, dtype=np.float64)
    e64 = rel(tr64["y"][::step], gold)
    print(f"shot {shot}: FP32 oracle vs golden {e32:.2e}, FP64 oracle vs golden {e64:.2e}")
    assert e32 < 1e-5 and e64 < 1e-5
    # the right-edge shots sit in the absorbing strip: their amplitude is lower
    amax = np.abs(gold).max()
    assert abs(np.abs(got).max() - amax) <= 1e-4 * amax


def test_su_header_layout(checker):
    """Golden SU headers: int32 slots 0, 9, 18-21; ns = nt at byte 114;
    dt field 0 because 60000 us exceeds 2^15 (solver/solver.h:38-108)."""
    _, d, _ = checker
    h = d["hdr_0"][0]
    i32 = h.view(np.int32)
    i16 = h.view(np.int16)
    assert i32[0] == 1
    assert i16[114 // 2] == 5000
    assert i16[116 // 2] == 0
    nonzero = set(np.nonzero(i32[:28])[0].tolist())
    assert nonzero <= {0, 9, 18, 19, 20, 21}

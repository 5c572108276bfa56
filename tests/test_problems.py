"""The input generators of inv.cu_amd/problems.py without a GPU.

* checker_gapped follows util/generate_checker.cu:4-70: with 4 x 4 squares,
  vs 3500 +- 400, rho 2600, vp 0 and sigma 3 it reproduces the reference's
  own example model (example/checker/model_true, held in
  tests/golden/checker_fixture.npz): x, z, vp and rho exactly, vs within
  1e-6 (the one-sided smoothing sums in another order than the reference's
  FP32 kernel).
* write_config writes a directory the command-line driver reads: the driver
  derives the grid from the point count and extent, the sources and stations
  land on the workload's grid indices through calcIdx, and the model files
  hold the workload's arrays.
"""
from pathlib import Path

import numpy as np
import pytest

import pyoracle
from conftest import inv

FIX = Path(__file__).resolve().parent / "golden" / "checker_fixture.npz"


def test_checker_gapped_reproduces_reference_example():
    d = np.load(FIX)
    x, z, vp, vs, rho = inv.problems.checker_gapped(4, 4, 201, 201, 2400.0, 2400.0, 0.0, 3500.0,
                                                    2600.0, 0.0, 400.0, 0.0, 3)
    assert np.array_equal(x, d["model_x"]) and np.array_equal(z, d["model_z"])
    assert np.array_equal(vp, d["model_vp"]) and np.array_equal(rho, d["model_rho"])
    e = np.abs(vs.astype(np.float64) - d["model_vs"]).max() / 3900.0
    assert e < 1e-6
    # a different sigma or square count is far off
    for args in ((4, 4, 2), (3, 3, 3)):
        *_, vs2, _ = inv.problems.checker_gapped(args[0], args[1], 201, 201, 2400.0, 2400.0,
                                                 0.0, 3500.0, 2600.0, 0.0, 400.0, 0.0, args[2])
        assert np.abs(vs2.astype(np.float64) - d["model_vs"]).max() > 10


def test_one_sided_gauss_matches_oracle():
    rng = np.random.default_rng(2)
    a = rng.standard_normal(37 * 29).astype(np.float32)
    got = inv.problems.one_sided_gauss(a, 37, 29, 4)
    ref = pyoracle.gauss(a, 37, 29, 4, dtype=np.float64)
    assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["c0", "c1", "c2"])
def test_write_config_round_trip(tmp_path, name):
    P = inv.problems.workload(name, nt=100)[0]
    root = inv.problems.write_config(P, tmp_path / name, mode=2, optimizer=1)
    cfg = {}
    for line in (root / "config.ini").read_text().splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            cfg[k.strip()] = v.split()[0]
    assert int(cfg["mode"]) == 2 and int(cfg["optimizer"]) == 1 and int(cfg["model_type"]) == 1
    assert int(cfg["nt"]) == 100 and int(cfg["sh"]) == P.sh and int(cfg["psv"]) == P.psv
    assert float(cfg["dt"]) == float(P.dt) and int(cfg["filter_sigma"]) == P.sigma
    x = pyoracle.read_bin(root / "model_true" / "proc000000_x.bin")
    z = pyoracle.read_bin(root / "model_true" / "proc000000_z.bin")
    xmax, zmax = float(x.max()), float(z.max())
    nx, nz, dx, dz = pyoracle.grid_dims(x.size, xmax, zmax)
    assert (nx, nz) == (P.nx, P.nz)
    src = pyoracle.read_dat(root / "sources.dat", 7)
    rec = pyoracle.read_dat(root / "stations.dat", 2)
    assert len(src) == P.nsrc and len(rec) == P.nrec
    assert [pyoracle.calc_idx(s[0], xmax, nx) for s in src] == list(P.src_ix)
    assert [pyoracle.calc_idx(s[1], zmax, nz) for s in src] == list(P.src_iz)
    assert [pyoracle.calc_idx(r[0], xmax, nx) for r in rec] == list(P.rec_ix)
    assert [pyoracle.calc_idx(r[1], zmax, nz) for r in rec] == list(P.rec_iz)
    types = P.src_type if P.src_type is not None else np.zeros(P.nsrc)
    assert np.array_equal(src[:, 2], types)
    # the wavelets the driver builds from the rows are the workload's
    stf = pyoracle.source(P.nt, P.dt, *src[0, 3:7])
    assert np.array_equal(stf[0], P.stf_x[0]) and np.array_equal(stf[1], P.stf_y[0])
    for m, arrs in (("model_true", (P.lam, P.mu, P.rho)), ("model_init", (P.lam0, P.mu0, P.rho0))):
        for c, a in zip(("lambda", "mu", "rho"), arrs):
            assert np.array_equal(pyoracle.read_bin(root / m / f"proc000000_{c}.bin"), a)

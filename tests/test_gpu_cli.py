"""The command-line driver (inv.cu_amd/inv.out, modes of inv.cu:6-82) on the
reference's checker example rebuilt from tests/golden/checker_fixture.npz.

The configuration below is example/checker/config.ini's (SH, nt 5000,
dt 0.06, absorbing width 20 / alpha 0.015 without the top edge, velocity
traces, envelope misfit, Gaussian filter sigma 5, L-BFGS with 5 updates);
model_init of the checker example is homogeneous (vp 0, vs 3500, rho 2600).

Checks:
  mode 3  SU traces of shots 0, 12, 24 against the reference's golden traces,
          SU headers byte for byte against the golden files' headers;
  mode 1  all sources at once equals the sum of the single-shot traces
          (linearity), wavefield snapshots independent of the export period;
  mode 5  the exported grid axes and vs against the model points;
  mode 2  kernels against the oracle's gradient (4 shots, nt 1000), and the
          observed-trace file path (trace_file = 1, velocity traces integrated
          to displacement) equal to the modelled one bit for bit;
  mode 0  one L-BFGS and one NLCG inversion lower the misfit, and the first
          model update is along minus the first filtered kernel.
"""
from __future__ import annotations

import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import pyoracle
from conftest import ROOT, inv, rel

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "checker_fixture.npz"
EXE = ROOT / "inv.cu_amd" / "inv.out"

CHECKER_CONFIG = {
    "mode": 0, "clean": 1, "solver": 0, "sh": 1, "psv": 0, "nt": 5000, "dt": 0.06,
    "model_type": 0, "trace_type": 0, "abs_bottom": 1, "abs_right": 1, "abs_top": 0,
    "abs_left": 1, "abs_width": 20, "abs_alpha": 0.015, "wfe": 200, "wae": 0, "sfe": 10,
    "misfit": 1, "filter": 0, "filter_sigma": 5, "trace_file": 0, "optimizer": 1,
    "inv_iteration": 5, "inv_cycle": 0, "inv_lambda": 0, "inv_mu": 1, "inv_rho": 0,
    "nlcg_type": 1, "nlcg_thresh": 0.0, "lbfgs_mem": 5, "lbfgs_thresh": 0.0, "ls_step": 10,
    "ls_step_max": 0.5, "ls_step_init": 0.05, "ls_thresh": 1.2,
}


def write_bin(path, a):
    a = np.ascontiguousarray(a, np.float32)
    with open(path, "wb") as f:
        f.write(np.int32(a.size).tobytes())
        f.write(a.tobytes())


def make_example(root: Path, nshots=None, **over):
    d = np.load(FIX)
    root.mkdir(parents=True, exist_ok=True)
    cfg = dict(CHECKER_CONFIG, **over)
    lines = ["[global]"] + [f"{k} = {v}" for k, v in cfg.items()]
    (root / "config.ini").write_text("\n".join(lines) + "\n")
    src = d["sources_txt"].tobytes().decode()
    if nshots is not None:
        rows = [l for l in src.splitlines() if l.strip() and not l.lstrip().startswith("#")]
        src = "# x | z | type | f0 | t0 | angle | amp\n" + "\n".join(rows[:nshots]) + "\n"
    (root / "sources.dat").write_text(src)
    (root / "stations.dat").write_bytes(d["stations_txt"].tobytes())
    n = d["model_x"].size
    for m, (vp, vs, rho) in {
        "model_true": (d["model_vp"], d["model_vs"], d["model_rho"]),
        "model_init": (np.zeros(n), np.full(n, 3500.0), np.full(n, 2600.0)),
    }.items():
        (root / m).mkdir(exist_ok=True)
        for c, a in (("x", d["model_x"]), ("z", d["model_z"]), ("vp", vp), ("vs", vs),
                     ("rho", rho)):
            write_bin(root / m / f"proc000000_{c}.bin", a)
    return root


def run_cli(cfgdir, *args, timeout=300):
    cmd = [str(EXE), str(cfgdir)] + [f"-{k}={v}" for k, v in args]
    print("running", " ".join(cmd), flush=True)
    assert EXE.exists(), "inv.out is not built (make -C inv.cu_amd)"
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    sys.stdout.write(p.stdout[-3000:])
    assert p.returncode == 0, p.stderr
    return p.stdout


def read_su(path, nt):
    raw = np.fromfile(path, np.uint8).reshape(-1, 240 + 4 * nt)
    return raw[:, :240].copy(), raw[:, 240:].copy().view(np.float32)


def read_bin(path):
    raw = Path(path).read_bytes()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    return np.frombuffer(raw[4:4 + 4 * n], np.float32).copy()


@pytest.fixture(scope="module")
def full(tmp_path_factory, gpu):
    return make_example(tmp_path_factory.mktemp("checker"))


@pytest.mark.parametrize("math,tol", [(1, 1e-5), (0, 1e-5)])
def test_mode3_traces_match_reference_golden(full, math, tol):
    run_cli(full, ("mode", 3), ("math", math))
    d = np.load(FIX)
    step = int(d["rec_step"])
    for shot in (0, 12, 24):
        hdr, data = read_su(full / "output" / f"vy_{shot:06d}.su", 5000)
        assert data.shape == (132, 5000)
        e = rel(data[::step], d[f"vy_{shot}"])
        print(f"math {math} shot {shot}: rel L2 vs golden {e:.2e}")
        assert e < tol
        assert np.array_equal(hdr[:2], d[f"hdr_{shot}"])


def test_mode1_linear_in_sources_and_snapshots(full):
    run_cli(full, ("mode", 3))
    single = sum(read_su(full / "output" / f"vy_{s:06d}.su", 5000)[1].astype(np.float64)
                 for s in range(25))
    run_cli(full, ("mode", 1), ("wfe", 1000))
    _, allsrc = read_su(full / "output" / "vy_000000.su", 5000)
    assert rel(allsrc, single) < 1e-4
    snaps = sorted(p.name for p in (full / "output").glob("proc*_vy.bin"))
    assert snaps == [f"proc{i * 1000:06d}_vy.bin" for i in range(1, 6)]
    last = read_bin(full / "output" / "proc005000_vy.bin")
    mid = read_bin(full / "output" / "proc002000_vy.bin")
    assert last.size == 201 * 201 and np.isfinite(last).all() and np.abs(mid).max() > 0
    run_cli(full, ("mode", 1), ("wfe", 2500))
    assert np.array_equal(read_bin(full / "output" / "proc005000_vy.bin"), last)


def test_mode5_model_export(full):
    run_cli(full, ("mode", 5))
    d = np.load(FIX)
    x, z = read_bin(full / "output" / "proc000000_x.bin"), read_bin(full / "output" / "proc000000_z.bin")
    i, j = np.meshgrid(np.arange(201), np.arange(201), indexing="ij")
    assert np.array_equal(x, (i * np.float32(2400.0)).ravel().astype(np.float32))
    assert np.array_equal(z, (j * np.float32(2400.0)).ravel().astype(np.float32))
    vs = read_bin(full / "output" / "proc000000_vs.bin")
    # the checker's points sit on the grid nodes: grid value = point value
    k = np.rint(d["model_x"] / 2400.0).astype(int) * 201 + np.rint(d["model_z"] / 2400.0).astype(int)
    ref = np.zeros(201 * 201, np.float32)
    ref[k] = d["model_vs"]
    assert rel(vs, ref) < 1e-6
    assert not (full / "output" / "proc000000_vp.bin").exists()  # inv_lambda = 0


def oracle_gradient(sigma, misfit_type, nshots, nt):
    d = np.load(FIX)
    x, z = d["model_x"], d["model_z"]
    xmax, zmax = float(x.max()), float(z.max())
    nx, nz, dx, dz = pyoracle.grid_dims(len(x), xmax, zmax)
    tmp = Path("/tmp") / "inv_cli_dat"
    tmp.mkdir(exist_ok=True)
    (tmp / "sources.dat").write_bytes(d["sources_txt"].tobytes())
    (tmp / "stations.dat").write_bytes(d["stations_txt"].tobytes())
    src = pyoracle.read_dat(tmp / "sources.dat", 7)[:nshots]
    recs = pyoracle.read_dat(tmp / "stations.dat", 2)
    dt = np.float32(0.06)
    ab = pyoracle.absbound(nx, nz, 20, 0.015, left=1, right=1, bottom=1, top=0)
    sx = [pyoracle.calc_idx(s[0], xmax, nx) for s in src]
    sz = [pyoracle.calc_idx(s[1], zmax, nz) for s in src]
    rx = [pyoracle.calc_idx(r[0], xmax, nx) for r in recs]
    rz = [pyoracle.calc_idx(r[1], zmax, nz) for r in recs]
    stf = [pyoracle.source(nt, dt, *s[3:7]) for s in src]
    n = len(x)

    def problem(vp, vs, rho):
        lam, mu, r = pyoracle.mesh2grid(nx, nz, dx, dz, xmax, zmax, x, z, vp, vs, rho)
        lam, mu = pyoracle.vps2lm(lam, mu, r)
        return pyoracle.Problem(nx, nz, dx, dz, dt, nt, lam, mu, r, ab, sx, sz,
                                np.stack([s[0] for s in stf]), np.stack([s[1] for s in stf]),
                                np.stack([s[2] for s in stf]), rx, rz, sh=1, psv=0, sfe=10,
                                adj=1, trace_type=0, inv_lambda=0, inv_mu=1, inv_rho=0)

    true = problem(d["model_vp"], d["model_vs"], d["model_rho"])
    obs = np.stack([true.forward(s)[0]["y"][None] for s in range(len(src))])
    init = problem(np.zeros(n), np.full(n, 3500.0), np.full(n, 2600.0))
    return init.gradient(obs, misfit_type=misfit_type, sigma=sigma)


@pytest.mark.parametrize("misfit", [0, 1])
def test_mode2_kernels_match_oracle(tmp_path, gpu, misfit):
    ex = make_example(tmp_path / "ex", nshots=4, nt=1000)
    run_cli(ex, ("mode", 2), ("misfit", misfit))
    km = read_bin(ex / "output" / "proc000000_k_mu.bin")
    assert not (ex / "output" / "proc000000_k_lambda.bin").exists()
    f, _, km_ref, _ = oracle_gradient(5, misfit, 4, 1000)
    e = rel(km, km_ref)
    print(f"misfit {misfit}: k_mu rel L2 vs oracle {e:.2e} (oracle misfit {f:.6e})")
    assert e < 1e-4


def test_mode2_trace_files_equal_modelled_obs(tmp_path, gpu):
    ex = make_example(tmp_path / "ex", nshots=4, nt=1000)
    run_cli(ex, ("mode", 3))
    (ex / "traces").mkdir()
    for s in range(4):
        (ex / "output" / f"vy_{s:06d}.su").rename(ex / "traces" / f"vy_{s:06d}.su")
    run_cli(ex, ("mode", 2))
    k_model = read_bin(ex / "output" / "proc000000_k_mu.bin")
    run_cli(ex, ("mode", 2), ("trace_file", 1))
    k_file = read_bin(ex / "output" / "proc000000_k_mu.bin")
    assert np.abs(k_model).max() > 0
    assert np.array_equal(k_file, k_model)


@pytest.mark.parametrize("optimizer", [1, 0])
def test_mode0_inversion(tmp_path, gpu, optimizer):
    ex = make_example(tmp_path / "ex", nshots=4, nt=1000)
    out = run_cli(ex, ("mode", 0), ("optimizer", optimizer), ("inv_iteration", 2), timeout=600)
    final = float(re.search(r"Final misfit: ([0-9.eE+-]+)", out).group(1))
    assert 0 < final < 1
    o = ex / "output"
    for name in ("proc000000_vs.bin", "proc000001_vs.bin", "proc000001_k_mu.bin",
                 "proc000002_vs.bin"):
        assert (o / name).exists(), name
    # first iteration: p = -g, m1 = m0 + alpha p  (mu = rho vs^2)
    rho = np.float64(2600.0)
    mu0 = rho * read_bin(o / "proc000000_vs.bin").astype(np.float64) ** 2
    mu1 = rho * read_bin(o / "proc000001_vs.bin").astype(np.float64) ** 2
    g = read_bin(o / "proc000001_k_mu.bin").astype(np.float64)
    dm = mu1 - mu0
    cos = -(dm @ g) / (np.linalg.norm(dm) * np.linalg.norm(g))
    print(f"optimizer {optimizer}: final misfit {final:.4f}, cos(dm, -g) {cos:.6f}")
    assert cos > 0.999


def test_mode4_device_query(tmp_path, gpu):
    """Mode 4 lists the devices (lib/query.h:23-65) and runs nothing else."""
    (tmp_path / "q").mkdir()
    (tmp_path / "q" / "config.ini").write_text("mode = 4\n")
    out = run_cli(tmp_path / "q")
    assert "HIP device(s) detected" in out and "gfx950" in out
    assert "Global memory" in out or "global memory" in out


def test_written_config_runs_the_library_gradient(tmp_path, gpu):
    """A configuration directory written by problems.write_config for the
    checkerboard workload (4 explosive shots, nt 400, sigma 5): mode 2
    exports the kernels the library's own gradient gives for the same
    problem, bit for bit (same batch, tile, math and graph settings)."""
    P = inv.problems.workload("c0", nt=400)[0]
    root = inv.problems.write_config(P, tmp_path / "c0", mode=2, batch=4)
    print(f"c0 written config, {P.nx}x{P.nz} nt={P.nt}", flush=True)
    run_cli(root)
    got = [read_bin(root / "output" / f"proc000000_k_{c}.bin") for c in ("lambda", "mu", "rho")]
    ctx = P.context(batch=4, model="true")
    ctx.forward(0, 4, snapshots=False)
    ctx.obs_from_traces(0, 4)
    P.load_model(ctx, "init")
    ctx.gradient(P.misfit, P.sigma)
    want = ctx.kernels_get()
    ctx.close()
    for c, a, b in zip(("lambda", "mu", "rho"), got, want):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b), f"k_{c}: rel L2 {rel(a, b):.3e}"


@pytest.mark.parametrize("fmt,ckpt,tune", [(3, 0, 0), (1, 100, 0), (0, 0, 1)],
                         ids=["i16", "f16-ckpt100", "f32-autotune"])
def test_written_config_snapshot_and_tuning_keys(tmp_path, gpu, fmt, ckpt, tune):
    """The driver's extension keys snap_format / ckpt (inv_set_snapshots) and
    autotune (inv_autotune at the first model load): mode 2 exports the
    kernels the library's gradient gives with the same storage, bit for bit
    (16-bit snapshots on the default tiles; FP32 snapshots give the same bits
    on any tile the autotuner picks)."""
    P = inv.problems.workload("c0", nt=400)[0]
    root = inv.problems.write_config(P, tmp_path / "c0", mode=2, batch=4, snap_format=fmt,
                                     ckpt=ckpt, autotune=tune)
    print(f"c0 written config, snapshots {fmt}, ckpt {ckpt}, autotune {tune}", flush=True)
    run_cli(root)
    got = [read_bin(root / "output" / f"proc000000_k_{c}.bin") for c in ("lambda", "mu", "rho")]
    ctx = P.context(batch=4, model="true")
    ctx.forward(0, 4, snapshots=False)
    ctx.obs_from_traces(0, 4)
    P.load_model(ctx, "init")
    ctx.set_snapshots(fmt, ckpt)
    ctx.gradient(P.misfit, P.sigma)
    want = ctx.kernels_get()
    ctx.close()
    for c, a, b in zip(("lambda", "mu", "rho"), got, want):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b), f"k_{c}: rel L2 {rel(a, b):.3e}"

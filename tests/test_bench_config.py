"""bench.py's workload set-up without a GPU: the shots per launch, shots per
step, snapshot format and checkpoint interval each workload runs with (every
16-bit snapshot resident where the device holds them), the overrides, and the
one-GPU guard."""
import importlib.util
import subprocess
import sys

from conftest import ROOT, inv


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _plan(args):
    b = _bench()
    a = b.parse(args)
    P, B, fmt, ckpt = inv.problems.workload(a.workload, nt=a.nt, sigma=a.sigma)
    wave = B if a.workload == "c4" else P.nsrc
    B, wave, fmt, ckpt = b.launch_plan(a, a.workload, P, B, fmt, ckpt, wave)
    return P, B, wave, fmt, ckpt, b.describe(a.workload, P, B, wave, fmt, ckpt)


def test_default_workload_is_the_large_psv_one():
    b = _bench()
    a = b.parse([])
    assert a.workload == "c4" and a.gpus == 1


def test_sh_workload_keeps_every_snapshot():
    P, B, wave, fmt, ckpt, desc = _plan(["--workload", "c2", "--nt", "200"])
    assert (P.nx, P.nz, P.nsrc, P.sh, P.psv) == (1474, 480, 64, 1, 0)
    assert fmt == inv.invhip.INV_SNAP_I16
    assert B == 64 and wave == 64 and ckpt == 0
    assert "64 per launch" in desc and "every snapshot resident" in desc


def test_large_psv_workload_steps_one_wave_of_four_shots():
    P, B, wave, fmt, ckpt, desc = _plan(["--workload", "c4", "--nt", "100"])
    assert (P.nx, P.nz, P.nsrc, P.nrec, P.misfit) == (4096, 2048, 256, 2048, 1)
    assert B == 4 and wave == 4 and ckpt == 0 and fmt == inv.invhip.INV_SNAP_I16
    # FP16 u and v of one shot at every 10th step must fit four times in HBM
    snap_bytes = 4 * 2 * P.nx * P.nz * (10000 // 10)
    assert 4 * snap_bytes < 280e9 < 5 * snap_bytes
    assert "4 of 256 shots per step, 4 per launch" in desc


def test_overrides_and_interval_that_does_not_divide_nt():
    _, B, wave, _, ckpt, desc = _plan(["--workload", "c4", "--nt", "100", "--batch", "16",
                                       "--ckpt", "50"])
    assert B == 16 and wave == 16 and ckpt == 50 and "50-step checkpoints" in desc
    _, B, _, _, ckpt, _ = _plan(["--workload", "c4", "--nt", "100", "--ckpt", "30"])
    assert ckpt == 100  # 30 does not divide nt: one segment
    _, B, _, fmt, _, _ = _plan(["--workload", "c2", "--nt", "200", "--batch", "500",
                                "--snap", "bf16"])
    assert B == 64 and fmt == inv.invhip.INV_SNAP_BF16  # clamped to the shots


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr

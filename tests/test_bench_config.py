"""bench.py's workload set-up without a GPU: the shots per launch and
checkpoint interval each workload runs with (every 16-bit snapshot resident
where the device holds them), the --batch / --ckpt overrides, and the
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
    P, B, fmt, ckpt, step_bytes, desc = b.workload(inv, a)
    return (P, fmt, step_bytes) + b.launch_plan(a, P, B, ckpt, desc)


def test_sh_workload_keeps_every_snapshot():
    P, fmt, step_bytes, B, ckpt, desc = _plan(["--workload", "c2", "--nt", "200"])
    assert (P.nx, P.nz, P.nsrc, P.sh, P.psv) == (1474, 480, 64, 1, 0)
    assert fmt == inv.invhip.INV_SNAP_F16 and step_bytes == 40
    assert B == 64 and ckpt == 0
    assert "64 per launch" in desc and "no checkpoints" in desc


def test_large_psv_workload_runs_four_shots_per_launch():
    P, fmt, step_bytes, B, ckpt, desc = _plan(["--workload", "c4", "--nt", "100"])
    assert (P.nx, P.nz, P.nsrc) == (4096, 2048, 16) and step_bytes == 68
    assert B == 4 and ckpt == 0
    # FP16 u and v of one shot at every 10th step must fit four times in HBM
    snap_bytes = 4 * 2 * P.nx * P.nz * (10000 // 10)
    assert 4 * snap_bytes < 280e9 < 5 * snap_bytes
    assert "4 per launch" in desc


def test_overrides_and_interval_that_does_not_divide_nt():
    _, _, _, B, ckpt, desc = _plan(["--workload", "c4", "--nt", "100", "--batch", "16",
                                    "--ckpt", "50"])
    assert B == 16 and ckpt == 50 and "50-step checkpoints" in desc
    _, _, _, B, ckpt, _ = _plan(["--workload", "c4", "--nt", "100", "--ckpt", "30"])
    assert ckpt == 100  # 30 does not divide nt: one segment
    _, _, _, B, ckpt, _ = _plan(["--workload", "c2", "--nt", "200", "--batch", "500"])
    assert B == 64  # clamped to the shots


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr

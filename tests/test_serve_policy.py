"""Host-only checks of the request server's admission and batching policy
(inv.cu_amd/csrc/serve_policy.h, compiled here with g++ from
tests/serve_policy_unit.cpp; no device call)."""
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_serve_policy(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    exe = tmp_path / "serve_policy_unit"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-o", str(exe),
                    str(ROOT / "tests" / "serve_policy_unit.cpp")], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(p.stdout)
    assert p.returncode == 0 and p.stdout.strip().endswith("ALL OK"), p.stdout + p.stderr

"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host-side code:
the CPU oracle (driven by tests/oracle_san.cpp over every entry point), the
C++ plugin layer (tests/host_unit.cpp over inv.cu_amd/host/inv.h) and the
request server's admission policy (tests/serve_policy_unit.cpp).  Built with
g++ on the host only; no device call is made.  A sanitizer report aborts the
program (-fno-sanitize-recover), so a clean run is exit status 0 and "ALL OK"."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=undefined"]


def _run(cmd, exe, args=(), env_extra=None, timeout=300):
    subprocess.run(cmd, check=True)
    env = dict(os.environ, UBSAN_OPTIONS="print_stacktrace=1")
    env.update(env_extra or {})
    p = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=timeout,
                       env=env)
    out = p.stdout + p.stderr
    assert p.returncode == 0 and p.stdout.strip().endswith("ALL OK"), out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]


@pytest.fixture(scope="module")
def gxx():
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    return "g++"


def test_oracle_sanitized(gxx, tmp_path):
    exe = tmp_path / "oracle_san"
    _run([gxx, "-std=c++17", *SAN, "-fopenmp", "-ffp-contract=off", "-o", str(exe),
          str(ROOT / "tests" / "oracle_san.cpp"), str(ROOT / "oracle" / "oracle.cpp")],
         exe, env_extra={"ASAN_OPTIONS": "detect_leaks=1", "OMP_NUM_THREADS": "2"})


def test_host_layer_sanitized(gxx, tmp_path):
    lib = ROOT / "inv.cu_amd"
    assert (lib / "libinv_hip.so").exists(), "libinv_hip.so is not built"
    exe = tmp_path / "host_unit"
    # leak checking is off: the HIP runtime that libinv_hip.so loads keeps
    # process-lifetime allocations
    _run([gxx, "-std=c++17", *SAN, "-Wall", "-o", str(exe),
          str(ROOT / "tests" / "host_unit.cpp"), f"-L{lib}", "-linv_hip",
          f"-Wl,-rpath,{lib}", "-Wl,-rpath,/opt/rocm/lib"],
         exe, args=[str(tmp_path / "cfg")], env_extra={"ASAN_OPTIONS": "detect_leaks=0"})


def test_serve_policy_sanitized(gxx, tmp_path):
    exe = tmp_path / "serve_policy_unit"
    _run([gxx, "-std=c++17", *SAN, "-Wall", "-Wextra", "-o", str(exe),
          str(ROOT / "tests" / "serve_policy_unit.cpp")],
         exe, env_extra={"ASAN_OPTIONS": "detect_leaks=1"})

"""Host-only unit checks of the C++ plugin layer (tests/host_unit.cpp): config
parsing, line-search rules, L-BFGS slot rotation.  Compiled here with g++
against libinv_hip.so's header; no device call is made."""
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_host_plugin_layer(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    exe = tmp_path / "host_unit"
    lib = ROOT / "inv.cu_amd"
    assert (lib / "libinv_hip.so").exists(), "libinv_hip.so is not built"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-o", str(exe),
                    str(ROOT / "tests" / "host_unit.cpp"), f"-L{lib}", "-linv_hip",
                    f"-Wl,-rpath,{lib}", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    p = subprocess.run([str(exe), str(tmp_path / "cfg")], capture_output=True, text=True,
                       timeout=60)
    print(p.stdout)
    assert p.returncode == 0 and p.stdout.strip().endswith("ALL OK"), p.stdout + p.stderr

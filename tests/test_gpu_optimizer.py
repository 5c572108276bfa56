"""The optimizers of the C++ plugin layer on the GPU (tests/opt_unit.cpp,
built with the library): NLCG directions against the Fletcher-Reeves and
Polak-Ribiere closed forms and the restart rules (optimizer/nlcg.h:8-52),
L-BFGS directions against the textbook two-loop recursion with the
reference's slot rotation (optimizer/lbfgs.h:18-91), and whole iterations
with the line search (optimizer/optimizer.h:360-467) lowering a quadratic
misfit every iteration."""
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_optimizers_on_a_quadratic(gpu, tmp_path):
    exe = ROOT / "tests" / "_bin" / "opt_unit"
    assert exe.exists(), "tests/_bin/opt_unit is not built (make -C inv.cu_amd)"
    print("optimizer checks on a 48x40 quadratic", flush=True)
    p = subprocess.run([str(exe), str(tmp_path / "opt")], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout)
    assert p.returncode == 0 and p.stdout.strip().endswith("ALL OK"), p.stdout + p.stderr

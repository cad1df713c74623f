"""The reverse-mode primitives of the kernels, compiled for the host: inverse
round trips and vector-Jacobian products against central differences
(tests/cpu/adjoint_check.cpp).  No GPU involved."""
import os
import subprocess
import tempfile

from conftest import ROOT


def test_adjoint_primitives_match_finite_differences():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "adjcheck")
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "accel-plus_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpu", "adjoint_check.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "all adjoint primitives ok" in out.stdout

"""The regrouped forward maps of the tracking kernel, compiled for the host,
against the same maps in the reference's order (tests/cpu/forward_check.cpp).
No GPU involved."""
import os
import subprocess
import tempfile

from conftest import ROOT


def test_regrouped_maps_match_reference_order():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "fwdcheck")
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "accel-plus_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpu", "forward_check.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "all regrouped maps ok" in out.stdout

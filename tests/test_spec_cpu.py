"""The lattice-specialized tracker's generator on the CPU
(tests/cpu/spec_check.cpp against libaccp.so): repeated cells found in
grouped element tables, refusals, and the generated kernel source compiled
for gfx950 by hiprtc for every group form of the lean tracker.  No GPU
involved."""
import os
import subprocess
import tempfile

from conftest import ROOT


def test_specialized_generator():
    pkg = os.path.join(ROOT, "accel-plus_amd")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "speccheck")
        subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I", os.path.join(pkg, "csrc"), os.path.join(ROOT, "tests", "cpu", "spec_check.cpp"),
                        "-L", pkg, "-laccp", f"-Wl,-rpath,{pkg}", "-Wl,-rpath,/opt/rocm/lib", "-o", exe],
                       check=True)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stdout
    assert "specialized tracker generator ok" in out.stdout

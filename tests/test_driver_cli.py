"""The command-file driver's parsing and error behaviour, without a GPU:
--parse-only reads the command file and the lattice file (the reference's
grammar) and stops before any device work; tracking without a device fails
with an error instead of falling back to the CPU."""
import os
import subprocess

import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from conftest import ROOT

RUN = os.path.join(ROOT, "accel-plus_amd", "accp-run")


def run(tmp_path, text, *flags):
    assert os.path.exists(RUN), "accp-run is not built (make -C accel-plus_amd)"
    (tmp_path / "cmd.txt").write_text(text)
    return subprocess.run([RUN, *flags, "cmd.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=60)


CMD = """# comments and continuation lines as the reference reads them
Lattice   % the first block
  file_name = da.lat, \\
      type=ERING
  set_flags: radiation=radQE, cavity=on, wake=off
  modify_element: type=Quad, famname=QF, field=K, delta=0.001
  correct_tune: type=Quad, famname=(QF QD) [0.27, 0.26] REPEAT
  export_element: type=Quad
End
Optics
  print_tune
End
Track
  Distribution: Zero NPar=4
  Monitor: interval=10, printflag=Stat2
  Do_Track: NTurn=100, radiation=off, RFCavity=on
End
DAMA
  TRACK_DA: xmin=0.001, xmax=0.02, ymax=0.01, nray=11, nstep=10, nturn=64
End
STOP
Lattice
  this block is after STOP and never read
End
"""


def test_parse_only_reads_every_block(tmp_path):
    el.write_lattice(lat.da_ring(0), str(tmp_path / "da.lat"))
    p = run(tmp_path, CMD, "--parse-only")
    assert p.returncode == 0, p.stderr
    out = p.stdout
    assert "Total elements: 322" in out
    assert "radiation turned on with quantum excitation" in out
    assert "1 cavity(ies) turned on" in out
    assert "EXPORT_ELEMENT is not carried by this backend" in p.stderr
    for block in ("OPTICS", "TRACK", "DAMA"):
        assert f"\n{block}\n" in out
    assert "#0.001 0.02 0.01 11 10 64 0" in out


def test_lattice_grammar_errors(tmp_path):
    (tmp_path / "bad.lat").write_text("D1 : Drift 1.0\nW1 : Wiggler 1 2 3\nRETURN\n")
    p = run(tmp_path, "Lattice\n file_name = bad.lat\nEnd\n", "--parse-only")
    assert p.returncode == 1 and "type Wiggler is not carried" in p.stderr
    p = run(tmp_path, "Track\nEnd\n", "--parse-only")
    assert p.returncode == 1 and "No LATTICE block" not in p.stdout and "first block must be LATTICE" in p.stdout


def test_tracking_without_a_device_is_an_error(tmp_path, lib):
    if lib.device_count() > 0:
        pytest.skip("a HIP device is visible: the driver would track")
    el.write_lattice(lat.fodo_ring(), str(tmp_path / "fodo.lat"))
    p = run(tmp_path, "Lattice\n file_name = fodo.lat\nEnd\nOptics\n print_tune\nEnd\n")
    assert p.returncode == 1 and "no HIP device" in p.stderr

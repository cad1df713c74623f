"""The command-file driver's parsing and error behaviour, without a GPU:
--parse-only reads the command file and the lattice file (the reference's
grammar) and stops before any device work; tracking without a device fails
with an error instead of falling back to the CPU."""
import os
import subprocess

import numpy as np
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
  print_lat_summary
End
Optics
  print_tune
  print_matrix: type=Quad, famname=QF, index=[1,2], line_style=oneline
  print_optics: type=Quad, famname=QD
End
Track
  Distribution: Zero NPar=4
  Monitor: interval=10, printflag=Stat2
  Do_Track: NTurn=100, radiation=off, RFCavity=on
  Transport: type=Quad, famname=QF, index=[1,3]
End
DAMA
  TRACK_DA: xmin=0.001, xmax=0.02, ymax=0.01, nray=11, nstep=10, nturn=64
  TRACK_DNUDJ: xmax=0.004, ymax=0.003, nturn=256
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
    assert "not carried" not in p.stderr
    for block in ("OPTICS", "TRACK", "DAMA"):
        assert f"\n{block}\n" in out
    assert "Quad QF 1 3 " in out  # the transport watch list as the reference echoes it
    assert "#0.004 0.003 256" in out  # TRACK_DNUDJ echo
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


def export_block(out):
    lines = out.splitlines()
    i = lines.index("$$$Export-Elements***********************")
    body = []
    for s in lines[i + 1:]:
        if s.startswith("***") or s in ("OPTICS", "TRACK", "DAMA", "LATTICE") or s.startswith("Total elements"):
            break
        body.append(s)
    return body


def test_export_is_lossless_and_reads_back(tmp_path):
    """EXPORT_ELEMENT writes the reference's lattice grammar at 17
    significant digits: read back, the ring is the same to the last bit."""
    ring = lat.da_ring(3, extra_quads=True)
    el.write_lattice(ring, str(tmp_path / "da.lat"))
    p = run(tmp_path, "Lattice\n file_name = da.lat\n export_element: type=RING\nEnd\n", "--parse-only")
    assert p.returncode == 0, p.stderr
    body = export_block(p.stdout)
    (tmp_path / "again.lat").write_text("\n".join(body) + "\nRETURN\n")
    q = run(tmp_path, "Lattice\n file_name = again.lat\n export_element: type=RING\nEnd\n", "--parse-only")
    assert q.returncode == 0, q.stderr
    assert export_block(q.stdout) == body
    # the numbers themselves: every K and multipole coefficient survives
    vals = [float(t) for s in body for t in s.replace(":", " ").split()[2:] if t.replace(".", "").replace("-", "")
            .replace("e", "").replace("+", "").isdigit()]
    ref = []
    for e in ring:
        ref.append(e.len)
        if e.type == el.QUAD:
            ref.append(e.k)
        if e.type == el.STRMPOLE:
            ref += list(e.polyB[:e.max_npole]) + list(e.polyA[:e.max_npole])
    assert set(ref) <= set(vals)


def test_rfmult_noise_is_not_read_from_a_lattice_file(tmp_path, reference):
    """The reference parses RFMult's noise sigmas and then builds the element
    with a constructor that clears them (accel.cpp:1065-1071, accel.h:376),
    while CrabCav passes its own to the constructor (accel.cpp:943-945).  The
    driver reads both files the same way: the exported RFMult carries zero
    sigmas, the CrabCav its own; the compiled reference, loading the same
    file, reports SigPhi 0 and 0.01."""
    from pyoracle import RefLine
    ring = [el.rfmult(0.1, 5e8, 0.2, [1e-4, 2e-3], [0.0, 1e-3], sigphi=0.01, sigdvv=0.02, name="RFM"),
            el.drift(0.5), el.crabcav(0.2, 1.5, 0.5, 1.5e9, 0.3, 3000.0, sigphi=0.01, sigdvv=0.02, name="CRAB")]
    path = tmp_path / "noise.lat"
    el.write_lattice(ring, str(path))
    p = run(tmp_path, "Lattice\n file_name = noise.lat\n export_element: type=RING\nEnd\n", "--parse-only")
    assert p.returncode == 0, p.stderr
    body = export_block(p.stdout)
    rfm = [float(t) for t in body[[i for i, s in enumerate(body) if "RFMult" in s][0]].split()[3:9]]
    crab = [float(t) for t in body[[i for i, s in enumerate(body) if "CrabCav" in s][0]].split()[3:11]]
    assert rfm[4:6] == [0.0, 0.0], rfm
    assert crab[6:8] == [0.01, 0.02], crab
    h = reference.lib.ref_load(str(path).encode())
    assert h
    line = RefLine(reference.lib, h)
    assert line.getfield(0, "SigPhi") == 0.0
    assert line.getfield(2, "SigPhi") == 0.01


def test_apply_setpt(tmp_path):
    """Set points (process.cpp:2273-2392): the first element must match, then
    one PolyB[2] value per SF member."""
    el.write_lattice(lat.da_ring(0), str(tmp_path / "da.lat"))
    vals = [20.0 + 0.125 * i for i in range(20)]
    (tmp_path / "sp.txt").write_text("$FirstElement: type=Drift, famname=D1\n"
                                     "$SetValue: type=StrMPole, famname=SF, PolyB[2]\n" +
                                     " ".join(repr(v) for v in vals) + "\n")
    p = run(tmp_path, "Lattice\n file_name = da.lat\n apply_setpt = sp.txt\n"
                      " export_element: type=StrMPole, famname=SF\nEnd\n", "--parse-only")
    assert p.returncode == 0, p.stderr
    assert "#first element matches" in p.stdout
    body = export_block(p.stdout)
    b2 = [float(body[3 * k + 1].split()[2]) for k in range(20)]
    assert b2 == vals
    (tmp_path / "bad.txt").write_text("$FirstElement: type=Quad, famname=QF\n")
    p = run(tmp_path, "Lattice\n file_name = da.lat\n apply_setpt = bad.txt\nEnd\n", "--parse-only")
    assert "#first element check: 0" in p.stdout


def test_gaussian_emittance_form(tmp_path):
    """Distribution: Gaussian with emit_x=, betx= (process.cpp:1619-1735) draws
    the same beam as sigx = sqrt(betx emit_x), sigxp = sqrt(emit_x / betx);
    betx <= 0 falls back to 1 m with the reference's message."""
    el.write_lattice(lat.fodo_ring(), str(tmp_path / "fodo.lat"))
    ex, bx, ey, by = 2.5e-9, 7.3, 4.0e-10, 2.9
    sig = [(ex * bx) ** 0.5, (ex / bx) ** 0.5, (ey * by) ** 0.5, (ey / by) ** 0.5]

    def beam(spec):
        p = run(tmp_path, f"""Lattice
  file_name = fodo.lat
End
Track
  Distribution: Gaussian NPar=50, {spec}, sigz=0.003, sigdpp=0.001, print
End
""", "--parse-only", "--precision", "17")
        assert p.returncode == 0, p.stderr
        lines = p.stdout.splitlines()
        i = lines.index("$$$Init-Distribution*************")
        return p.stdout, [[float(t) for t in s.split()] for s in lines[i + 1:i + 51]]

    _, a = beam(f"emit_x={ex!r}, betx={bx!r}, emit_y={ey!r}, bety={by!r}")
    _, b = beam("sigx={!r}, sigxp={!r}, sigy={!r}, sigyp={!r}".format(*sig))
    assert len(a) == 50 and a == b
    out, c = beam(f"emit_x={ex!r}, betx=-1, sigy=1e-4, sigyp=1e-5")
    assert "#betx=1.0 m is assumed" in out
    _, d = beam(f"sigx={ex ** 0.5!r}, sigxp={ex ** 0.5!r}, sigy=1e-4, sigyp=1e-5")
    assert c == d


def test_reset_first_and_print_elem_value(tmp_path):
    """RESET_FIRST rotates the line (aline.cpp:557-605; process.cpp:254-340)
    and PRINT_ELEM_VALUE prints the selected fields of the family in the new
    order (process.cpp:344-480).  With type= alone the reference's parser
    consumes the first name to print; that is reproduced."""
    el.write_lattice(lat.fodo_ring(), str(tmp_path / "fodo.lat"))
    p = run(tmp_path, """Lattice
  file_name = fodo.lat
  reset_first: type=Quad, famname=QD, index=2
  print_elem_value: type=Quad, famname=QF, Index, Length, K
  print_elem_value: type=Quad Index Length K
  reset_first: famname=SD
  print_elem_value: type=StrMPole, famname=SD, Index, PolyB[2], PolyA[2]
End
""", "--parse-only")
    assert p.returncode == 0, p.stderr
    blocks = p.stdout.split("$$$Print-Element_Values***********************\n")[1:]
    rows = [b.split("******")[0].strip("\n").split("\n") for b in blocks]
    # QD #2 sits at 19: QF members 1, 13, ... move to (i - 19) mod 240
    qf = sorted((i - 19) % 240 for i in range(1, 240, 12))
    assert rows[0] == [f"{i}\t0.3\t1.2\t" for i in qf]
    assert len(rows[1]) == 40 and set(rows[1]) == {"0.3\t1.2\t", "0.3\t-1.1\t"}
    # famname alone: every SD met while walking the line moves the first SD
    # to the front again (20 messages, one rotation)
    assert p.stdout.count("First element moved to: type=StrMPole, famname=SD, index= 1") == 20
    assert rows[2] == [f"{i}\t-25\t0\t" for i in range(0, 240, 12)]


def test_wigtable_lattice_reads_and_exports_losslessly(tmp_path):
    """A WigTable record in the reference grammar (accel.cpp:1298-1335:
    header, x grid, y grid, x-kick rows, blank line, y-kick rows) reads in,
    and EXPORT_ELEMENT writes it back (AWigTable::serialize layout) so that
    every grid value and kick survives to the last bit."""
    from tests.golden.make_golden import wigtable_cases
    cases = [e for _, e in wigtable_cases()]
    ring = [el.drift(1.0)] + cases + [el.quad(0.3, 1.2)]
    el.write_lattice(ring, str(tmp_path / "wt.lat"))
    p = run(tmp_path, "Lattice\n file_name = wt.lat\n export_element: type=RING\nEnd\n", "--parse-only")
    assert p.returncode == 0, p.stderr
    assert f"Total elements: {len(ring)}" in p.stdout
    body = export_block(p.stdout)
    (tmp_path / "again.lat").write_text("\n".join(body) + "\nRETURN\n")
    q = run(tmp_path, "Lattice\n file_name = again.lat\n export_element: type=RING\nEnd\n", "--parse-only")
    assert q.returncode == 0, q.stderr
    assert export_block(q.stdout) == body
    # the tables themselves, row by row
    k = 0
    for e in cases:
        while "WigTable" not in body[k]:
            k += 1
        head = body[k].split()
        assert [int(head[4]), int(head[5]), int(head[6])] == [e.nslice, e.wt_mx, e.wt_my]
        assert float(head[3]) == e.len
        xl, yl, xk, yk = el.wigtable_arrays(e)
        rows = [np.array([float(t) for t in s.split()]) for s in body[k + 1:k + 3 + 2 * e.wt_my + 1]]
        assert np.array_equal(rows[0], xl) and np.array_equal(rows[1], yl)
        assert np.array_equal(np.array(rows[2:2 + e.wt_my]), xk)
        assert rows[2 + e.wt_my].size == 0
        assert np.array_equal(np.array(rows[3 + e.wt_my:]), yk)
        k += 1

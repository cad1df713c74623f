"""The command-file driver (accp-run) on the GPU.

Command files in the reference's grammar (src/process.cpp) run through the
driver, which prints the reference's blocks; numbers are read back at 17
digits and compared with the oracle on the same lattice and beam, and with
the compiled reference's fixtures where those exist."""
import os
import subprocess

import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from conftest import ROOT, golden
from helpers import assert_coords_close

pytestmark = pytest.mark.gpu

RUN = os.path.join(ROOT, "accel-plus_amd", "accp-run")


def run(tmp_path, text, check=True, flags=()):
    """The driver's output lines; with check=False the finished process."""
    (tmp_path / "cmd.txt").write_text(text)
    print("accp-run", *flags, tmp_path / "cmd.txt", flush=True)
    p = subprocess.run([RUN, "--precision", "17", *flags, "cmd.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    if not check:
        return p
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr
    return p.stdout.splitlines()


def after(lines, head, n):
    i = next(k for k, s in enumerate(lines) if s.startswith(head))
    return lines[i + 1:i + 1 + n]


def numbers(s):
    """The numeric tokens of one printed line."""
    out = []
    for t in s.replace("[", " ").replace("]", " ").replace(",", " ").split():
        try:
            out.append(float(t))
        except ValueError:
            pass
    return out


def test_ering_tune_and_tracking(tmp_path, oracle):
    ring = lat.da_ring(0)
    el.write_lattice(ring, str(tmp_path / "da.lat"))
    beam = lat.gaussian_beam(200, seed=4)
    el.write_plain_beam(beam, str(tmp_path / "beam.txt"))
    out = run(tmp_path, """Lattice
  file_name = da.lat, type=ERING
End
Optics
  print_tune
End
Track
  Distribution: Load beam.txt, format=plain
  Monitor: interval=5, printflag=Stat1
  Do_Track: NTurn=20, radiation=off, RFCavity=on, PrintFinal=Full
End
STOP
""")
    arr = el.to_array(ring)
    phis = oracle.setrfphase(arr)
    got = float(next(s for s in out if s.startswith("RF cavities phase")).split("=")[1].split()[0])
    assert abs(got - phis) <= 1e-12 * abs(phis)
    tune = numbers(next(s for s in out if s.startswith("Tunes:")))[:2]
    assert np.allclose(tune, oracle.tune(arr), rtol=0, atol=1e-12)
    # tracking with the cavity switched on by Do_Track
    for i in range(len(arr)):
        if arr[i].type == el.CAVITY:
            arr[i].on = 1
    stats = [numbers(s) for s in after(out, "$$$Track-Data: Stat1", 6)[2:6]]
    r = beam.copy()
    for k, rec in enumerate(stats):
        ref = oracle.beam_stat(r)
        assert rec[0] == 5 * k and rec[1] == ref[1]
        assert np.allclose(rec[2:14], ref[2:14], rtol=1e-10, atol=1e-18), (k, rec, ref)
        r = oracle.track(arr, r, 5)
    final = np.array([numbers(s)[1:7] for s in after(out, "$$$Track-Data: Full export", 201)[1:]])
    assert final.shape == (200, 6)
    assert_coords_close(final, r, 1e-10, what="20 turns through the driver")


def test_specialized_tracking(tmp_path, oracle):
    """Do_Track with --specialized (the kernel compiled for the lattice's
    structure): the final coordinates against the oracle, as without it, and
    the monitor statistics against the default kernel's."""
    ring = lat.da_ring(1)
    el.write_lattice(ring, str(tmp_path / "da.lat"))
    beam = lat.gaussian_beam(300, seed=6)
    el.write_plain_beam(beam, str(tmp_path / "beam.txt"))
    cmd = """Lattice
  file_name = da.lat, type=ERING
End
Track
  Distribution: Load beam.txt, format=plain
  Monitor: interval=4, printflag=Stat1
  Do_Track: NTurn=12, radiation=off, RFCavity=on, PrintFinal=Full
End
STOP
"""
    spec = run(tmp_path, cmd, flags=("--specialized",))
    walk = run(tmp_path, cmd)
    arr = el.to_array(ring)
    oracle.setrfphase(arr)  # Do_Track's RFCavity=on sets the synchronous phase, as the reference does
    for i in range(len(arr)):
        if arr[i].type == el.CAVITY:
            arr[i].on = 1
    final = np.array([numbers(s)[1:7] for s in after(spec, "$$$Track-Data: Full export", 301)[1:]])
    assert_coords_close(final, oracle.track(arr, beam, 12), 1e-10, what="12 turns, --specialized")
    st_s = np.array([numbers(s) for s in after(spec, "$$$Track-Data: Stat1", 5)[2:5]])
    st_w = np.array([numbers(s) for s in after(walk, "$$$Track-Data: Stat1", 5)[2:5]])
    assert np.array_equal(st_s[:, :2], st_w[:, :2])
    np.testing.assert_allclose(st_s[:, 2:], st_w[:, 2:], rtol=1e-10, atol=1e-18)


def test_track_with_wake_and_monitor(tmp_path, oracle):
    """Do_Track on a ring holding a TransWake switched on by set_flags, with
    monitor statistics (the turn splits at the wake): the Stat1 records and
    the final coordinates against the oracle."""
    ring = lat.fodo_ring() + [el.transwake(0.0, 2.0, 2.0e5, 1.5, 5.0, 2.0e5, 2.0, 3.0, on=1)]
    el.write_lattice(ring, str(tmp_path / "tw.lat"))
    beam = lat.gaussian_beam(400, seed=23)
    el.write_plain_beam(beam, str(tmp_path / "beam.txt"))
    out = run(tmp_path, """Lattice
  file_name = tw.lat
  set_flags: wake=on
End
Track
  Distribution: Load beam.txt, format=plain
  Monitor: interval=4, printflag=Stat1
  Do_Track: NTurn=8, radiation=off, RFCavity=off, PrintFinal=Full
End
STOP
""")
    arr = el.to_array(ring)
    stats = [numbers(s) for s in after(out, "$$$Track-Data: Stat1", 4)[2:4]]
    r = beam.copy()
    for k, rec in enumerate(stats):
        ref = oracle.beam_stat(r)
        assert rec[0] == 4 * k and rec[1] == ref[1]
        assert np.allclose(rec[2:14], ref[2:14], rtol=1e-9, atol=1e-18), (k, rec, ref)
        r = oracle.track(arr, r, 4, turn0=4 * k)
    final = np.array([numbers(s)[1:7] for s in after(out, "$$$Track-Data: Full export", 401)[1:]])
    assert_coords_close(final, r, 1e-9, what="8 turns with a transverse wake through the driver")


def test_track_da_with_loss_turns(tmp_path):
    """TRACK_DA sets radiation (classical) and the cavity on, like the
    reference's DAMA block; with ERING phasing this is the trackda fixture's
    setup."""
    g = golden("trackda.npz")
    el.write_lattice(lat.da_ring(0), str(tmp_path / "da.lat"))
    out = run(tmp_path, """Lattice
  file_name = da.lat, type=ERING
End
DAMA
  TRACK_DA: xmin=0.001, xmax=0.02, ymax=0.01, nray=11, nstep=10, nturn=64, print=loss_turn
End
""")
    loss = np.array([numbers(s) for s in after(out, "$$$DA-Loss", 110)])
    assert loss.shape == (110, 3) and np.all((loss[:, 2] >= 0) & (loss[:, 2] < 64))
    prof = np.array([numbers(s) for s in after(out, "$$$DA-profile", 11)])
    step = (0.02 - 0.001) / 9
    moved = np.abs(np.hypot(prof[:, 0], prof[:, 1] * 2) - np.hypot(g["rx"], g["ry"] * 2))
    assert np.all(moved <= step * (1 + 1e-9)), moved
    area = numbers(next(s for s in out if s.startswith("$$$DA-profile")).split("=")[1])[0]
    if not moved.any():
        assert abs(area - g["area"][0] * 1e6) <= 1e-9 * abs(g["area"][0] * 1e6)


def test_correct_tune_and_chrom(tmp_path):
    g = golden("optics.npz")
    el.write_lattice(lat.fodo_ring(), str(tmp_path / "fodo.lat"))
    out = run(tmp_path, """Lattice
  file_name = fodo.lat
  correct_tune: type=Quad, famname=(QF QD) [0.28, 0.24] method=fd
  correct_chrom: type=StrMPole, famname=(SF SD) [2.0, 1.0] REPEAT
  correct_tune: type=Quad, famname=(QF QD) [0.27, 0.25] REPEAT method=adjoint
End
""")
    tunes = [numbers(s) for s in out if s.startswith("corrected tunes:")]
    chroms = [numbers(s) for s in out if s.startswith("corrected chromaticities:")]
    # method=fd is the reference's own step (finite-difference responses)
    assert np.allclose(tunes[0], g["corrected_tune"], rtol=0, atol=1e-9)
    # adjoint responses (the default): two Newton steps reach the targets
    assert np.allclose(chroms[-1], [2.0, 1.0], rtol=0, atol=1e-5)
    assert np.allclose(tunes[-1], [0.27, 0.25], rtol=0, atol=1e-8)
    assert not any(s.startswith("note:") for s in out)


def test_correction_falls_back_to_finite_differences(tmp_path):
    """A crab cavity switched on with phase noise has no reverse map:
    without METHOD the correction notes it and takes the reference's
    forward difference, with METHOD=ADJOINT it is an error.  The cavity
    has no voltage, so the FODO ring's reference step still applies."""
    g = golden("optics.npz")
    ring = lat.fodo_ring() + [el.crabcav(0.0, 0.0, 0.0, 1.5e9, 0.0, 3000.0, sigphi=1e-3)]
    el.write_lattice(ring, str(tmp_path / "crab.lat"))
    out = run(tmp_path, """Lattice
  file_name = crab.lat
  set_flags: crabcav=on
  correct_tune: type=Quad, famname=(QF QD) [0.28, 0.24]
End
""")
    assert any(s.startswith("note: finite-difference responses") for s in out)
    tunes = [numbers(s) for s in out if s.startswith("corrected tunes:")]
    assert np.allclose(tunes[0], g["corrected_tune"], rtol=0, atol=1e-9)
    p = run(tmp_path, """Lattice
  file_name = crab.lat
  set_flags: crabcav=on
  correct_tune: type=Quad, famname=(QF QD) [0.28, 0.24] method=adjoint
End
""", check=False)
    assert p.returncode != 0 and "no reverse map" in p.stderr


def test_optics_and_momentum_aperture_commands(tmp_path, oracle):
    """PRINT_OPTICS on the QFs at dp = 0 against the oracle's calctwiss
    (bit-exact with the reference's); TRACK_MA against the reference's
    momentum apertures (optics2.npz) within one step."""
    g = golden("optics2.npz")
    el.write_lattice(lat.fodo_ring(), str(tmp_path / "fodo.lat"))
    el.write_lattice(lat.da_ring(0), str(tmp_path / "da.lat"))
    out = run(tmp_path, """Lattice
  file_name = fodo.lat
End
Optics
  print_optics: type=Quad, famname=QF
End
Lattice
  file_name = da.lat, type=ERING
End
DAMA
  TRACK_MA: type=Quad, famname=QF, nmoni=4, dppmin=0.0005, dppmax=0.01, nstep=19, nturn=64
End
""")
    qf = lat.family(lat.fodo_ring(), "QF")
    rows = np.array([numbers(s) for s in after(out, "$$$Optics-Twiss", len(qf) + 1)[1:]])
    assert np.array_equal(rows[:, 0], qf)
    ref = oracle.calctwiss(lat.fodo_ring(), qf, 0.0)  # {s, psix, psiy, betx, alfx, bety, alfy}
    mine = rows[:, [1, 6, 7, 2, 4, 3, 5]]
    assert np.allclose(mine, ref, rtol=1e-10, atol=1e-12)
    ma = np.array([numbers(s) for s in after(out, "$$$MA", 4)])
    step = (0.01 - 0.0005) / 19
    assert np.all(np.abs(ma[:, 0] - g["ma_plus"]) <= step * (1 + 1e-9))
    assert np.all(np.abs(ma[:, 1] - g["ma_minus"]) <= step * (1 + 1e-9))


def test_accel_export_reload_is_lossless(tmp_path):
    """PrintFinal=Full writes 17 significant digits, and Distribution Load
    format=accel-export reads them back (process.cpp:1499-1545): 20 turns,
    dump, reload, 5 more turns equals 25 turns in one run, bit for bit."""
    el.write_lattice(lat.fodo_ring(), str(tmp_path / "fodo.lat"))
    el.write_plain_beam(lat.gaussian_beam(64, seed=9), str(tmp_path / "beam.txt"))
    first = run(tmp_path, """Lattice
  file_name = fodo.lat
End
Track
  Distribution: Load beam.txt
  Do_Track: NTurn=20, PrintFinal=Full
End
""")
    (tmp_path / "dump.txt").write_text("\n".join(first) + "\n")
    cont = run(tmp_path, """Lattice
  file_name = fodo.lat
End
Track
  Distribution: Load dump.txt, format=accel-export
  Do_Track: NTurn=5, PrintFinal=Full
End
""")
    once = run(tmp_path, """Lattice
  file_name = fodo.lat
End
Track
  Distribution: Load beam.txt
  Do_Track: NTurn=25, PrintFinal=Full
End
""")
    a = after(cont, "$$$Track-Data: Full export", 65)[1:]
    b = after(once, "$$$Track-Data: Full export", 65)[1:]
    # rows are "turn<TAB>x<TAB>...": the coordinates must agree to the last digit
    assert len(a) == 64 and [r.split("\t")[1:] for r in a] == [r.split("\t")[1:] for r in b]


def test_gaussian_emittance_uses_ring_twiss(tmp_path, oracle):
    """Distribution: Gaussian emit_x= without betx takes betx (and bety) from
    the periodic Twiss functions at the ring start (process.cpp:1640-1650):
    the beam equals the sigx/sigxp form with the oracle's calctwiss."""
    el.write_lattice(lat.fodo_ring(), str(tmp_path / "fodo.lat"))
    tw = oracle.calctwiss0(lat.fodo_ring())
    ex, ey = 3.0e-9, 5.0e-10
    bx, by = float(tw[3]), float(tw[5])
    sig = [(ex * bx) ** 0.5, (ex / bx) ** 0.5, (ey * by) ** 0.5, (ey / by) ** 0.5]

    def beam(spec):
        out = run(tmp_path, f"""Lattice
  file_name = fodo.lat
End
Track
  Distribution: Gaussian NPar=40, {spec}, print
End
""")
        i = out.index("$$$Init-Distribution*************")
        return np.array([numbers(s) for s in out[i + 1:i + 41]])

    a = beam(f"emit_x={ex!r}, emit_y={ey!r}")
    b = beam("sigx={!r}, sigxp={!r}, sigy={!r}, sigyp={!r}".format(*sig))
    assert a.shape == (40, 6) and np.abs(a).max() > 0
    assert np.allclose(a, b, rtol=1e-12, atol=0)


def test_lattice_summary(tmp_path, oracle):
    """PRINT_LAT_SUMMARY (process.cpp:482-512) against the oracle: tunes and
    chromaticities with radiation off, the 6D closed orbit (radiation and
    cavities then stay on, as getco6 leaves them) and the ring-start Twiss
    functions of that state."""
    ring = lat.da_ring(0)
    el.write_lattice(ring, str(tmp_path / "da.lat"))
    out = run(tmp_path, """Lattice
  file_name = da.lat, type=ERING
  print_lat_summary
End
""")
    arr = el.to_array(ring)
    oracle.setrfphase(arr)
    for i in range(len(arr)):
        if arr[i].type == el.CAVITY:
            arr[i].on = 0
        if arr[i].type == el.BENDMPOLE:
            arr[i].rad = 0
    head = after(out, "$$$Lattice-Summary", 9)
    assert head[0].startswith(f"Total {len(arr)} elements, total length = ")
    assert float(head[0].split("=")[1].split()[0]) == pytest.approx(sum(e.len for e in ring), rel=1e-15)
    assert np.allclose(numbers(head[2])[:2], oracle.tune(arr), rtol=0, atol=1e-12)
    assert np.allclose(numbers(head[3])[:2], oracle.chrom(arr), rtol=1e-6, atol=1e-6)
    # (I - M) has a condition number of a few hundred (slow synchrotron
    # motion): 1e-10 of the orbit, as for getco6 itself
    co = oracle.getco6(arr)
    assert np.all(np.abs(np.array(numbers(head[5])) - co) <= 1e-10 * np.abs(co).max())
    tw = oracle.calctwiss0(arr)
    assert np.allclose(numbers(head[7]), [tw[3], tw[5], tw[4], tw[6]], rtol=1e-9, atol=1e-12)


def test_track_dnudj(tmp_path):
    """TRACK_DNUDJ (process.cpp:2587-2640) on the DA ring: the twelve
    amplitude/tune lines trackdnudexy prints and the four coefficients,
    against the reference's (naff.npz) to the NAFF search tolerance."""
    g = golden("naff.npz")
    el.write_lattice(lat.da_ring(2), str(tmp_path / "da.lat"))
    out = run(tmp_path, """Lattice
  file_name = da.lat
End
DAMA
  TRACK_DNUDJ: xmax=0.004, ymax=0.003, nturn=256
End
""")
    i = out.index("$$$dtuning coefficients")
    rows = np.array([numbers(s) for s in out[i - 12:i]])
    assert rows.shape == (12, 3)
    d = np.array(numbers(out[i + 1]))
    assert np.allclose(d, g["dnudexy_da"], rtol=1e-4, atol=0), (d, g["dnudexy_da"])
    # amplitudes grow in each scan; tunes stay within 0.1 of the linear ones
    assert np.all(np.diff(rows[:6, 0]) > 0) and np.all(np.diff(rows[6:, 0]) > 0)


def test_transport_to_watch_points(tmp_path, oracle):
    """TRANSPORT (process.cpp:2108-2214): the beam passes [0, w0) and then
    [w(k-1), wk) for QF members 1, 3 and 5 ([1,3,5], 1-based); the printed
    coordinates at each watch point against the oracle over the same
    ranges (the transported copy, where the reference prints its
    untransported input)."""
    ring = lat.fodo_ring()
    el.write_lattice(ring, str(tmp_path / "fodo.lat"))
    beam = lat.gaussian_beam(10, seed=8)
    el.write_plain_beam(beam, str(tmp_path / "beam.txt"))
    out = run(tmp_path, """Lattice
  file_name = fodo.lat
End
Track
  Distribution: Load beam.txt
  Transport: type=Quad, famname=QF, index=[1,3,5]
End
""")
    qf = lat.family(ring, "QF")
    watch = [qf[0], qf[2], qf[4]]
    r = beam.copy()
    start = 0
    arr = el.to_array(ring)
    for w in watch:
        i = out.index(next(s for s in out if s.startswith(f"type=Quad, famname=QF, ring index={w},")))
        spos = float(out[i].split("spos=")[1])
        assert spos == pytest.approx(sum(e.len for e in ring[:w]), rel=1e-14)
        got = np.array([numbers(s) for s in out[i + 1:i + 11]])
        assert np.all(got[:, 0] == w)
        r = oracle.track(arr, r, 1, st=start, ed=w)
        assert_coords_close(got[:, 1:], r, 1e-12, what=f"transport to {w}")
        start = w


def test_print_matrix_selections_and_styles(tmp_path, oracle):
    """PRINT_MATRIX (process.cpp:1228-1325, print_m66_to_line :1397-1425):
    the one-turn matrix in "oneline" style and the matrices from the ring
    start to QF members 1, 2 and 3 ("[1,2,3]") in "4by4" style, against the
    oracle's getm66 over the same element ranges."""
    ring = lat.fodo_ring()
    el.write_lattice(ring, str(tmp_path / "fodo.lat"))
    out = run(tmp_path, """Lattice
  file_name = fodo.lat
End
Optics
  print_matrix: type=RING, line_style=oneline
  print_matrix: type=Quad, famname=QF, index=[1,2,3], line_style=4by4
End
""")
    i = out.index("$$$OPTICS-MATRIX*******************ONE-TURN")
    one = np.array(numbers(out[i + 2]))
    assert one.shape == (36,)
    assert np.allclose(one.reshape(6, 6), oracle.getm66(ring), rtol=1e-9, atol=1e-12)
    qf = lat.family(ring, "QF")
    j = next(k for k, s in enumerate(out) if s.startswith("$$$OPTICS-MATRIX, dpp=0"))
    assert out[j].endswith(" 3 points")
    for n, e in enumerate(qf[:3]):
        head = out[j + 1 + 5 * n]
        assert head.startswith(f"type=Quad, famname=QF, ring index={e},")
        m4 = np.array([numbers(s) for s in out[j + 2 + 5 * n:j + 6 + 5 * n]])
        assert m4.shape == (4, 4)
        assert np.allclose(m4, oracle.getm66(ring, st=0, ed=e)[:4, :4], rtol=1e-9, atol=1e-12), e

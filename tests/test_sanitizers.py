"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the
CPU (no GPU involved):

* the oracle (oracle/accp_oracle.c) built with -fsanitize=address,undefined
  and driven through the entry points the tests use (tracking with
  LongWake and every element type, getm66, tunes, chromaticities, twiss,
  closed orbit, DA, tracktune, naff, beam statistics), by
  tests/cpu/oracle_sanitize_main.c;
* the command-file driver (accel-plus_amd/host/accp_run.cpp) built the same
  way and run with --parse-only over command files with every block, a
  lattice of every element type with EXPORT_ELEMENT and MODIFY_ELEMENT, and
  malformed inputs: a missing file, an unknown type, MaxNPole and WigTable
  sizes out of range must end in an error message, and a truncated
  multipole record (which the reference's stream reads leave at zero) in
  neither an error nor a fault.

Any report aborts the program (-fno-sanitize-recover=all, leak detection
on), so a clean exit is the check.  ThreadSanitizer is not used: the only
threads on these paths are OpenMP's in the oracle, whose runtime is not
instrumented, so it reports the runtime's own synchronisation."""
import os
import subprocess

import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from conftest import ROOT

SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def dump(ring, path):
    """accp_elem records, each WigTable's table after its record."""
    with open(path, "wb") as f:
        for e in el.to_array(ring):
            f.write(bytes(e))
            if e.type == el.WIGTABLE:
                k = e.wt_mx + e.wt_my + 2 * e.wt_mx * e.wt_my
                f.write(np.ctypeslib.as_array(e.wt_table, (k,)).tobytes())


def test_oracle_under_asan_ubsan(tmp_path):
    from test_integration_binding import every_type
    exe = str(tmp_path / "oracle_sanitize")
    subprocess.run(["gcc", *SAN, "-fopenmp", "-std=gnu99", "-Wall", "-I", os.path.join(ROOT, "oracle"), "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "oracle", "accp_oracle.c"),
                    os.path.join(ROOT, "tests", "cpu", "oracle_sanitize_main.c"), "-lm", "-o", exe], check=True)
    lat.gaussian_beam(200, seed=4).tofile(str(tmp_path / "beam.bin"))
    rings = {"da_wake": lat.da_ring(0) + [el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.15, 0.15)],
             "every_type": every_type()}
    for name, ring in rings.items():
        dump(ring, str(tmp_path / f"{name}.bin"))
        p = subprocess.run([exe, f"{name}.bin", "beam.bin"], cwd=tmp_path, capture_output=True, text=True,
                           env=ENV, timeout=600)
        print(name, p.stdout.strip())
        assert p.returncode == 0 and "checksum" in p.stdout, p.stderr[-3000:]


def test_driver_parser_under_asan_ubsan(tmp_path):
    from test_driver_cli import CMD
    from test_integration_binding import every_type
    exe = str(tmp_path / "accp-run-san")
    lib = os.path.join(ROOT, "accel-plus_amd")
    subprocess.run(["g++", *SAN, "-std=c++17", "-Wall", os.path.join(lib, "host", "accp_run.cpp"), "-L", lib,
                    "-laccp", "-Wl,-rpath," + lib, "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe], check=True)
    el.write_lattice(lat.da_ring(0), str(tmp_path / "da.lat"))
    el.write_lattice(every_type(), str(tmp_path / "every.lat"))
    good = {
        "blocks": CMD,
        "every": ("Lattice\n file_name = every.lat\n export_element: type=RING\n"
                  " modify_element: type=StrMPole, famname=SF, field=PolynomB, index=1, value=2.5\n"
                  " modify_element: type=Quad, famname=QF, field=K, delta=1e-3\n"
                  " export_element: type=StrMPole, famname=SF\n print_lat_summary\nEnd\n"),
    }
    (tmp_path / "bad_type.lat").write_text("Q1 : Solenoid 0.5 1.0\nRETURN\n")
    (tmp_path / "truncated.lat").write_text("SF : StrMPole 0.2 3\n 0 0\n")
    (tmp_path / "npole.lat").write_text("SF : StrMPole 0.2 4000000\n 0 0 1\n 0 0 0\nRETURN\n")
    (tmp_path / "wigtab.lat").write_text("WT : WigTable 0.5 4 500 3\n 0 1\n 0 1 2\nRETURN\n")
    bad = {name: f"Lattice\n file_name = {name}.lat\nEnd\n"
           for name in ("missing", "bad_type", "truncated", "npole", "wigtab")}
    for name, text in list(good.items()) + list(bad.items()):
        (tmp_path / f"{name}.txt").write_text(text)
        p = subprocess.run([exe, "--parse-only", f"{name}.txt"], cwd=tmp_path, capture_output=True, text=True,
                           env=ENV, timeout=300)
        print(name, p.returncode, p.stderr.strip()[-200:])
        assert "Sanitizer" not in p.stderr and "runtime error" not in p.stderr, (name, p.stderr[-3000:])
        if name in good or name == "truncated":
            assert p.returncode == 0, (name, p.stderr[-2000:])
        else:
            assert p.returncode == 1 and "accp-run:" in p.stderr, (name, p.returncode, p.stderr[-2000:])

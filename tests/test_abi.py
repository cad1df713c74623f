"""The C ABI library on the CPU: it loads, exports every entry point the
header declares, its record layout matches the C compiler's, and it reports
errors instead of crashing (no GPU is needed for any of this)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import pyaccp
import pyaccp.abi as abi
import pyaccp.elements as el
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "accp.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(accp_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(lib):
    names = header_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", lib.path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (accp_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(abi.EXPORTS) == set(names)


def test_abi_version(lib):
    assert lib.lib.accp_abi_version() == 3  # accp_elem gained the WigTable grid sizes and table pointer


def test_elem_layout_matches_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "accp.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(accp_elem), offsetof(accp_elem, len), offsetof(accp_elem, wt_table),
         offsetof(accp_elem, polyB), offsetof(accp_elem, wake_rs), offsetof(accp_elem, famname),
         sizeof(accp_track_opts));
  printf("%zu %zu\n", sizeof(accp_tuning), sizeof(accp_param));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        a, b = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")[:2]
    vals = [int(v) for v in a.split()]
    E = el.Elem
    assert vals == [ctypes.sizeof(E), E.len.offset, E.wt_table.offset, E.polyB.offset, E.wake_rs.offset, E.famname.offset,
                    ctypes.sizeof(abi.TrackOpts)]
    assert [int(v) for v in b.split()] == [ctypes.sizeof(abi.Tuning), ctypes.sizeof(abi.Param)]


def test_bad_arguments_report_errors(lib):
    L = lib.lib
    h = ctypes.c_void_p()
    assert L.accp_lattice_build(None, 0, ctypes.byref(h)) == -1
    assert b"bad arguments" in L.accp_last_error()
    assert L.accp_track(None, None, 0, 1, 0, -1, None) == -1
    assert L.accp_batch_replay(None, None) == -5
    assert L.accp_lattice_free(None) == 0
    # NAFF: power-of-two lengths 4..4096 and an ordered window, checked
    # before any device work
    sig = (ctypes.c_double * 24)()
    out = (ctypes.c_double * 2)()
    assert L.accp_naff(sig, 2, 12, ctypes.c_double(0.0), ctypes.c_double(0.5), out) == -1
    assert b"power-of-two length" in L.accp_last_error()
    assert L.accp_naff(sig, 1, 8, ctypes.c_double(0.3), ctypes.c_double(0.2), out) == -1
    assert b"window" in L.accp_last_error()
    assert L.accp_tracktune(None, 100, ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0),
                            ctypes.c_double(0.5), ctypes.c_double(0), ctypes.c_double(0.5), out) == -1


def test_no_device_is_an_error_not_a_fallback(lib):
    if lib.device_count() > 0:
        pytest.skip("a GPU is visible; covered by the gpu tests")
    with pytest.raises(pyaccp.AccpError, match="no HIP device"):
        lib.lattice([el.drift(1.0)])
    with pytest.raises(pyaccp.AccpError, match="no HIP device"):
        lib.naff([[1.0, 0.0, -1.0, 0.0]])


def test_missing_library_raises():
    with pytest.raises(pyaccp.AccpError, match="not built"):
        abi.Library(os.path.join(ROOT, "does-not-exist.so"))

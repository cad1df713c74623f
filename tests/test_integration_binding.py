"""The reference-side binding of INTEGRATION.md, compiled and run on the CPU
against the compiled reference (oracle/_ref).  No GPU involved.

* The binding's aline_accp.cpp, extracted from INTEGRATION.md, compiles
  against the reference's own headers (src/lib/accel/accel.h).
* Linked with the reference's reader, its fill() turns every element type
  the backend carries, read from a lattice file, into the same accp_elem
  fields the Python elements that wrote the file hold.
* EXPORT_ELEMENT output of the driver (accp-run --parse-only) is read back
  by the reference's reader into the same records: the export is in the
  reference's grammar and loses nothing (17 significant digits)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from conftest import ROOT

REFLIB = "/root/reference/src/lib/accel"
REFSO = os.path.join(ROOT, "oracle", "_ref", "libaccel_ref.so")
RUN = os.path.join(ROOT, "accel-plus_amd", "accp-run")
LIBDIR = os.path.join(ROOT, "accel-plus_amd")

pytestmark = pytest.mark.skipif(not (os.path.isdir(REFLIB) and os.path.exists(REFSO)),
                                reason="needs the reference tree and oracle/_ref")


def binding_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    i = text.index("// src/lib/accel/aline_accp.cpp")
    return text[i:text.index("```", i)]


@pytest.fixture(scope="module")
def binding(tmp_path_factory):
    d = tmp_path_factory.mktemp("binding")
    (d / "aline_accp.cpp").write_text(binding_source())
    exe = str(d / "binding_main")
    subprocess.run(["g++", "-std=c++11", "-O1", "-fopenmp", "-I", REFLIB, "-I", os.path.join(ROOT, "include"),
                    "-I", str(d), os.path.join(ROOT, "tests", "cpu", "binding_main.cpp"), "-o", exe,
                    REFSO, "-L", LIBDIR, "-laccp", "-Wl,-rpath," + LIBDIR + ":" + os.path.dirname(REFSO)],
                   check=True)
    return d, exe


def test_binding_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "aline_accp.cpp"
    src.write_text(binding_source())
    subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-I", REFLIB, "-I",
                    os.path.join(ROOT, "include"), str(src)], check=True)


def every_type():
    """One element of every type the backend and the reference share, with
    what the lattice file carries (BendMPole's energy and the slice counts
    of BendMPole and StrMPole stay at the reference's defaults: its files
    hold neither)."""
    rng = np.random.default_rng(3)
    m = np.eye(6) + 0.01 * rng.standard_normal((6, 6))
    xl, yl = np.linspace(-0.02, 0.02, 5), np.linspace(-0.01, 0.01, 3)
    X, Y = np.meshgrid(xl, yl)
    return [
        el.drift(0.75, "D1"), el.quad(0.3, 1.234567890123, "QF"), el.corrector(0.1, 1.5e-4, -2.5e-4, "HC"),
        el.aperture(0.0, (-0.025, 0.02), (-0.012, 0.011), "AP"), el.monitor("BPM"),
        el.bend(1.2, 0.0785398163397, 0.01, 0.02, -0.15, "BL"),
        el.bendmpole(1.5, 2 * np.pi / 40, 0.012, 0.013, polyB=[0.0, -0.1, 2.5, 7.0]),
        el.strmpole(0.2, [0.0, 0.0, 20.5, 0.0, 1e3], polyA=[0.0, 1e-3, 0.0, 0.5, 0.0], name="SF"),
        el.strmpole(0.0, [1e-5, 2e-4, 3.0], polyA=[0.0, 0.0, 0.1], name="TK"),
        el.cavity(0.0, 1.5, 5.62e8, 0.3, 3000.0, name="RF"), el.matrix(0.4, m.ravel(), "MT"),
        el.syncrad(0.2, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0, name="SR"),
        el.crabcav(0.3, 1.5, 0.5, 1.5e9, 0.3, 3000.0, sigphi=1e-3, sigdvv=2e-3, name="CC"),
        el.rfmult(0.1, 5e8, 0.2, [1e-4, 2e-3, 0.5], [0.0, 1e-3, 0.2], energy=3000.0, name="RM"),
        el.idealwig(2.0, 20, 0.1, 0.5, 1, name="WG"),
        el.wigtable(0.5, 4, xl, yl, 1e-4 * np.sin(50 * X) * np.cosh(40 * Y), 1e-4 * np.cos(30 * X) * Y, name="WT"),
        el.transwake(0.0, 2.0, 1e3, 1.2, 5.0, 2e3, 1.5, 7.0, name="TW"),
    ]


FIELDS = {
    el.DRIFT: [], el.QUAD: ["k"], el.CORRECTOR: ["kickx", "kicky"], el.APERTURE: ["xlim", "ylim"],
    el.MONITOR: [], el.BEND: ["theta", "k", "e1", "e2"],
    el.BENDMPOLE: ["theta", "e1", "e2", "energy", "nslice", "max_npole", "polyB", "rad"],
    el.STRMPOLE: ["nslice", "max_npole", "polyB", "polyA"],
    el.CAVITY: ["volt", "freq", "phis", "energy", "on"], el.MATRIX: ["mat"],
    el.SYNCRAD: ["energy", "u0", "jx", "jy", "emitx", "emity", "sigdpp", "betax", "betay"],
    el.CRABCAV: ["devoltx", "devolty", "freq", "phis", "energy", "sigphi", "sigdvv", "on"],
    el.RFMULT: ["freq", "phis", "energy", "max_npole", "polyB", "polyA"],
    el.IDEALWIG: ["nslice", "wig_lambda", "wig_h", "wig_type", "rad"],
    el.WIGTABLE: ["nslice", "wt_mx", "wt_my"],
    el.TRANSWAKE: ["energy", "wake_charge", "tw_w", "tw_fr", "tw_q", "on"],
}


def records(exe, path, rad, n):
    out = str(path) + ".bin"
    p = subprocess.run([exe, str(path), out, str(rad)], capture_output=True, text=True)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = open(out, "rb").read()
    recs, tables, off = [], [], 0
    for _ in range(n):
        e = el.Elem.from_buffer_copy(raw[off:off + ctypes.sizeof(el.Elem)])
        off += ctypes.sizeof(el.Elem)
        t = None
        if e.type == el.WIGTABLE:
            k = e.wt_mx + e.wt_my + 2 * e.wt_mx * e.wt_my
            t = np.frombuffer(raw[off:off + 8 * k], dtype=np.float64)
            off += 8 * k
        recs.append(e)
        tables.append(t)
    assert off == len(raw)
    return recs, tables


def value(e, f):
    v = getattr(e, f)
    if f in ("polyB", "polyA"):
        return list(v)[:e.max_npole]
    return list(v) if hasattr(v, "__len__") and not isinstance(v, bytes) else v


def assert_same(recs, tables, ring, rad, what):
    assert len(recs) == len(ring), what
    for r, t, e in zip(recs, tables, ring):
        assert r.type == e.type and r.name == e.name and r.len == e.len, (what, e, r)
        for f in FIELDS[e.type]:
            want = value(e, f)
            if f == "rad":
                want = rad
            if f == "on":
                want = 1
            assert value(r, f) == want, (what, e, f, value(r, f), want)
        if e.type == el.WIGTABLE:
            k = e.wt_mx + e.wt_my + 2 * e.wt_mx * e.wt_my
            assert np.array_equal(t, np.ctypeslib.as_array(e.wt_table, (k,))), what


def test_binding_fill_reads_every_element_type(binding, tmp_path):
    d, exe = binding
    ring = every_type()
    path = tmp_path / "every.lat"
    el.write_lattice(ring, str(path))
    for rad in (0, 1):
        recs, tables = records(exe, path, rad, len(ring))
        assert_same(recs, tables, ring, rad, f"rad {rad}")


def test_export_element_is_read_by_the_reference(binding, tmp_path):
    """accp-run's EXPORT_ELEMENT of a lattice (a DA ring with its error
    kicks, and one element of every type), written to a file and loaded by
    the reference's reader: the same records as the original file's."""
    from test_driver_cli import export_block
    d, exe = binding
    for name, ring in (("da", lat.da_ring(2, extra_quads=True)), ("every", every_type())):
        src = tmp_path / f"{name}.lat"
        el.write_lattice(ring, str(src))
        (tmp_path / "cmd.txt").write_text(f"Lattice\n file_name = {name}.lat\n export_element: type=RING\nEnd\n")
        p = subprocess.run([RUN, "--parse-only", "cmd.txt"], cwd=tmp_path, capture_output=True, text=True,
                           timeout=60)
        assert p.returncode == 0, p.stderr
        exported = tmp_path / f"{name}_export.lat"
        exported.write_text("\n".join(export_block(p.stdout)) + "\nRETURN\n")
        a, ta = records(exe, src, 0, len(ring))
        b, tb = records(exe, exported, 0, len(ring))
        for x, y, u, v in zip(a, b, ta, tb):
            assert all(value(x, f) == value(y, f) for f in FIELDS[x.type] + ["type", "len", "famname"]), (name, x, y)
            if u is not None:
                assert np.array_equal(u, v)
        assert_same(b, tb, ring, 0, f"export of {name}")

"""TEST INFRASTRUCTURE ONLY: ctypes views of the CPU checker.

`Oracle` wraps liboracle.so (the plain-C restatement in accp_oracle.c);
`Reference` wraps _ref/libaccel_ref.so (the reference's own sources plus
ref_harness.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
import ctypes
import os
import tempfile
from ctypes import POINTER, c_char_p, c_double, c_int, c_uint, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DP = POINTER(c_double)
IP = POINTER(c_int)


def _dp(a):
    return a.ctypes.data_as(DP)


def _ip(a):
    return a.ctypes.data_as(IP)


class Oracle:
    def __init__(self, path=None):
        import pyaccp.elements as el

        self.el = el
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        L = ctypes.CDLL(path)
        E = POINTER(el.Elem)
        L.orc_line_pass.argtypes = [E, c_int, DP, c_int, c_int, c_int, c_int, c_uint64, c_uint64, c_uint64]
        L.orc_line_pass_omp.argtypes = [E, c_int, DP, c_int, c_int, c_int, c_uint64, c_uint64, c_uint64]
        L.orc_line_pass_omp.restype = c_int
        L.orc_elem_pass.argtypes = [E, DP, c_int, c_void_p]
        L.orc_getm66.argtypes = [E, c_int, c_int, c_int, c_double, c_double, DP]
        L.orc_calctune.argtypes = [E, c_int, c_double, c_double, DP, DP]
        L.orc_calcchrom.argtypes = [E, c_int, c_double, c_double, DP, DP]
        L.orc_calcresp.argtypes = [E, c_int, c_int, IP, c_int, IP, c_int, DP]
        L.orc_correct.argtypes = [E, c_int, c_int, IP, c_int, IP, c_int, DP, DP]
        L.orc_setrfphase.argtypes = [E, c_int]
        L.orc_setrfphase.restype = c_double
        L.orc_trackda.argtypes = [E, c_int, c_double, c_double, c_double, c_int, c_int, c_int, DP, DP, c_int]
        L.orc_trackda.restype = c_double
        L.orc_beam_stat.argtypes = [DP, c_int, c_double, DP]
        L.orc_beam_stat.restype = c_int
        L.orc_philox4x32_10.argtypes = [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)]
        L.orc_gauss.argtypes = [c_uint64, c_uint64, c_uint64, c_uint32, c_uint32]
        L.orc_gauss.restype = c_double
        L.orc_wake_w.argtypes = [c_double, c_double, c_double, c_double]
        L.orc_wake_w.restype = c_double
        L.orc_wake_pairwise.argtypes = [E, DP, c_int, DP]
        L.orc_getco4.argtypes = [E, c_int, c_double, DP]
        L.orc_getco6.argtypes = [E, c_int, DP]
        L.orc_calctwiss0.argtypes = [E, c_int, c_double, DP]
        L.orc_calctwiss.argtypes = [E, c_int, IP, c_int, c_double, DP]
        L.orc_calcchrom2.argtypes = [E, c_int, DP]
        L.orc_trackma.argtypes = [E, c_int, c_double, c_double, c_int, c_int, IP, c_int, DP, DP]
        L.orc_trackma.restype = c_double
        L.orc_naff.argtypes = [c_int, DP, c_double, c_double]
        L.orc_naff.restype = c_double
        L.orc_fprojection.argtypes = [c_double, c_int, DP]
        L.orc_fprojection.restype = c_double
        L.orc_tracktune.argtypes = [E, c_int, c_uint, c_double, c_double, c_double, c_double, c_double, c_double,
                                    DP]
        L.orc_trackdiffusion.argtypes = [E, c_int, c_int, c_int, DP, DP, DP, DP, DP, c_double, c_double,
                                         c_double, c_double]
        L.orc_trackdiffusion.restype = c_double
        L.orc_trackdnudexy.argtypes = [E, c_int, c_uint, c_double, c_double, DP]
        for nm in ("orc_quad6", "orc_drift_kzero", "orc_fastdrift"):
            getattr(L, nm).argtypes = [DP, c_double] + ([c_double] if nm == "orc_quad6" else [])
        self.lib = L

    def _arr(self, elems):
        return elems if isinstance(elems, ctypes.Array) else self.el.to_array(elems)

    def track(self, elems, r, nturn=1, st=0, ed=-1, seed=0, turn0=0, pid0=0):
        a = self._arr(elems)
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        n = r.size // 6
        self.lib.orc_line_pass(a, len(a), _dp(r), n, st, ed, nturn, seed, turn0, pid0)
        return r.reshape(-1, 6)

    def track_omp(self, elems, r, nturn, nthreads, seed=0, turn0=0, pid0=0):
        a = self._arr(elems)
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        rc = self.lib.orc_line_pass_omp(a, len(a), _dp(r), r.size // 6, nturn, nthreads, seed, turn0, pid0)
        if rc != 0:
            raise ValueError("collective element in an OpenMP pass")
        return r.reshape(-1, 6)

    def getm66(self, elems, dpp=0.0, delta=1e-8, st=0, ed=-1):
        a = self._arr(elems)
        m = np.zeros(36)
        self.lib.orc_getm66(a, len(a), st, ed, dpp, delta, _dp(m))
        return m.reshape(6, 6)

    def tune(self, elems, dpp=0.0, delta=1e-8):
        a = self._arr(elems)
        x, y = c_double(), c_double()
        self.lib.orc_calctune(a, len(a), dpp, delta, ctypes.byref(x), ctypes.byref(y))
        return x.value, y.value

    def chrom(self, elems, dp=0.0, delta=1e-8):
        a = self._arr(elems)
        x, y = c_double(), c_double()
        self.lib.orc_calcchrom(a, len(a), dp, delta, ctypes.byref(x), ctypes.byref(y))
        return x.value, y.value

    def resp(self, elems_arr, kind, fam1, fam2):
        f1 = np.asarray(fam1, dtype=np.int32)
        f2 = np.asarray(fam2, dtype=np.int32)
        out = np.zeros(4)
        self.lib.orc_calcresp(elems_arr, len(elems_arr), kind, _ip(f1), len(f1), _ip(f2), len(f2), _dp(out))
        return out.reshape(2, 2)

    def correct(self, elems_arr, kind, fam1, fam2, tx, ty):
        f1 = np.asarray(fam1, dtype=np.int32)
        f2 = np.asarray(fam2, dtype=np.int32)
        x, y = c_double(tx), c_double(ty)
        self.lib.orc_correct(elems_arr, len(elems_arr), kind, _ip(f1), len(f1), _ip(f2), len(f2),
                             ctypes.byref(x), ctypes.byref(y))
        return x.value, y.value

    def setrfphase(self, elems_arr):
        return self.lib.orc_setrfphase(elems_arr, len(elems_arr))

    def trackda(self, elems, xmax, xmin, ymax, nturn, nray, nstep, negweight=False):
        a = self._arr(elems)
        rx = np.zeros(51)
        ry = np.zeros(51)
        s = self.lib.orc_trackda(a, len(a), xmax, xmin, ymax, nturn, nray, nstep, _dp(rx), _dp(ry), int(negweight))
        nr = min(max(nray, 3), 51)
        return s, rx[:nr], ry[:nr]

    def beam_stat(self, r, turn=0.0):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(20)
        self.lib.orc_beam_stat(_dp(r), r.size // 6, turn, _dp(out))
        return out

    def getco4(self, elems_arr, dp=0.0):
        """Modifies the flags of elems_arr (radiation and cavities off)."""
        rc = np.zeros(4)
        self.lib.orc_getco4(elems_arr, len(elems_arr), dp, _dp(rc))
        return rc

    def getco6(self, elems_arr):
        """Modifies the flags of elems_arr (radiation and cavities on)."""
        rc = np.zeros(6)
        self.lib.orc_getco6(elems_arr, len(elems_arr), _dp(rc))
        return rc

    def calctwiss0(self, elems, dp=0.0):
        a = self._arr(elems)
        tw = np.zeros(7)
        self.lib.orc_calctwiss0(a, len(a), dp, _dp(tw))
        return tw

    def calctwiss(self, elems, idx, dp=0.0):
        a = self._arr(elems)
        ix = np.ascontiguousarray(idx, dtype=np.int32)
        tw = np.zeros((len(ix), 7))
        self.lib.orc_calctwiss(a, len(a), _ip(ix), len(ix), dp, _dp(tw))
        return tw

    def calcchrom2(self, elems):
        a = self._arr(elems)
        c = np.zeros(4)
        self.lib.orc_calcchrom2(a, len(a), _dp(c))
        return c

    def trackma(self, elems, dppmax, dppmin, nturn, nstep, locs):
        a = self._arr(elems)
        lc = np.ascontiguousarray(locs, dtype=np.int32)
        p, m = np.zeros(len(lc)), np.zeros(len(lc))
        avg = self.lib.orc_trackma(a, len(a), dppmax, dppmin, nturn, nstep, _ip(lc), len(lc), _dp(p), _dp(m))
        return avg, p, m

    def naff(self, data, xlow=0.0, xhi=0.5):
        d = np.ascontiguousarray(data, dtype=np.float64)
        return self.lib.orc_naff(len(d), _dp(d), xlow, xhi)

    def fprojection(self, nv, data):
        d = np.ascontiguousarray(data, dtype=np.float64)
        return self.lib.orc_fprojection(nv, len(d), _dp(d))

    def tracktune(self, elems, nturn, x0, y0, xlow=0.0, xhi=0.5, ylow=0.0, yhi=0.5):
        a = self._arr(elems)
        t = np.zeros(2)
        self.lib.orc_tracktune(a, len(a), nturn, x0, y0, xlow, xhi, ylow, yhi, _dp(t))
        return t

    def trackdiffusion(self, elems, nturn, x0, y0, xlow=0.0, xhi=0.5, ylow=0.0, yhi=0.5):
        a = self._arr(elems)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        nux, nuy, dif = np.zeros(len(x0)), np.zeros(len(x0)), np.zeros(len(x0))
        tot = self.lib.orc_trackdiffusion(a, len(a), nturn, len(x0), _dp(x0), _dp(y0), _dp(nux), _dp(nuy),
                                          _dp(dif), xlow, xhi, ylow, yhi)
        return tot, nux, nuy, dif

    def trackdnudexy(self, elems, nturn, xm, ym):
        a = self._arr(elems)
        d = np.zeros(4)
        self.lib.orc_trackdnudexy(a, len(a), nturn, xm, ym, _dp(d))
        return d

    def philox(self, ctr, key):
        c = (c_uint32 * 4)(*ctr)
        k = (c_uint32 * 2)(*key)
        o = (c_uint32 * 4)()
        self.lib.orc_philox4x32_10(c, k, o)
        return list(o)

    def gauss(self, seed, pid, turn, eidx, slc):
        return self.lib.orc_gauss(seed, pid, turn, eidx, slc)

    def wake_w(self, rs, q, fr, tau):
        return self.lib.orc_wake_w(rs, q, fr, tau)

    def wake_pairwise(self, e, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(r.size // 6)
        self.lib.orc_wake_pairwise(ctypes.byref(e), _dp(r), r.size // 6, _dp(out))
        return out


class Reference:
    """The compiled reference (oracle/_ref/libaccel_ref.so)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_ref", "libaccel_ref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        L = ctypes.CDLL(path)
        L.ref_load.argtypes = [c_char_p]
        L.ref_load.restype = c_void_p
        L.ref_free.argtypes = [c_void_p]
        L.ref_nelem.argtypes = [c_void_p]
        L.ref_set_rad.argtypes = [c_void_p, c_int]
        L.ref_set_cavity.argtypes = [c_void_p, c_int]
        L.ref_set_crab.argtypes = [c_void_p, c_int]
        L.ref_set_wake.argtypes = [c_void_p, c_int]
        L.ref_set_threads.argtypes = [c_int]
        L.ref_pass.argtypes = [c_void_p, DP, c_int, c_int]
        L.ref_pass_range.argtypes = [c_void_p, DP, c_int, c_int, c_int]
        L.ref_elem_pass.argtypes = [c_void_p, c_int, DP, c_int]
        L.ref_setrfphase.argtypes = [c_void_p]
        L.ref_setrfphase.restype = c_double
        L.ref_getm66.argtypes = [c_void_p, c_double, c_double, DP]
        L.ref_calctune.argtypes = [c_void_p, DP, DP]
        L.ref_calcchrom.argtypes = [c_void_p, DP, DP]
        L.ref_correct.argtypes = [c_void_p, c_int, c_char_p, c_char_p, c_char_p, DP, DP]
        L.ref_calcresp.argtypes = [c_void_p, c_int, c_char_p, c_char_p, c_char_p, DP]
        L.ref_trackda.argtypes = [c_void_p, c_double, c_double, c_double, c_int, c_int, c_int, DP, DP, c_int]
        L.ref_trackda.restype = c_double
        L.ref_beam_stat.argtypes = [DP, c_int, DP, DP, DP]
        L.ref_beam_stat.restype = c_int
        L.ref_setfield.argtypes = [c_void_p, c_int, c_char_p, c_double]
        L.ref_setfield_int.argtypes = [c_void_p, c_int, c_char_p, c_int]
        L.ref_set_mpole.argtypes = [c_void_p, c_int, c_char_p, c_int, c_double]
        L.ref_getfield.argtypes = [c_void_p, c_int, c_char_p]
        L.ref_getfield.restype = c_double
        L.ref_get_mpole.argtypes = [c_void_p, c_int, c_char_p, c_int]
        L.ref_get_mpole.restype = c_double
        IPT = POINTER(c_int)
        L.ref_getco4.argtypes = [c_void_p, c_double, DP]
        L.ref_getco6.argtypes = [c_void_p, DP]
        L.ref_calctwiss0.argtypes = [c_void_p, c_double, DP]
        L.ref_calctwiss.argtypes = [c_void_p, IPT, c_int, c_double, DP]
        L.ref_calcchrom2.argtypes = [c_void_p, DP]
        L.ref_trackma.argtypes = [c_void_p, c_double, c_double, c_int, c_int, IPT, c_int, DP, DP]
        L.ref_trackma.restype = c_double
        L.ref_naff.argtypes = [c_int, DP, c_double, c_double]
        L.ref_naff.restype = c_double
        L.ref_tracktune.argtypes = [c_void_p, c_uint, c_double, c_double, c_double, c_double, c_double, c_double,
                                    DP]
        L.ref_trackdiffusion.argtypes = [c_void_p, c_int, c_int, DP, DP, DP, DP, DP, c_double, c_double,
                                         c_double, c_double]
        L.ref_trackdiffusion.restype = c_double
        L.ref_trackdnudexy.argtypes = [c_void_p, c_uint, c_double, c_double, DP]
        L.ref_pass_crn.argtypes = [c_void_p, DP, c_int, c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.ref_pass_crn.restype = c_int
        self.lib = L

    def naff(self, data, xlow=0.0, xhi=0.5):
        d = np.ascontiguousarray(data, dtype=np.float64)
        return self.lib.ref_naff(len(d), _dp(d), xlow, xhi)

    def load(self, elems):
        """Write `elems` as a lattice file, load it, then apply the state the
        file cannot carry (nslice, BendMPole energy, radiation, cavity flags)."""
        import pyaccp.elements as el

        fd, path = tempfile.mkstemp(suffix=".lat")
        os.close(fd)
        try:
            el.write_lattice(elems, path)
            h = self.lib.ref_load(path.encode())
        finally:
            os.unlink(path)
        if not h:
            raise RuntimeError("reference failed to load lattice")
        line = RefLine(self.lib, h)
        if self.lib.ref_nelem(h) != len(elems):
            raise RuntimeError(f"reference parsed {self.lib.ref_nelem(h)} of {len(elems)} elements")
        for i, e in enumerate(elems):
            if e.type in (el.BENDMPOLE, el.STRMPOLE) and e.nslice != 10:
                self.lib.ref_setfield_int(h, i, b"NSlice", e.nslice)
            if e.type == el.BENDMPOLE and e.energy != el.DEFAULT_E0:
                self.lib.ref_setfield(h, i, b"energy", e.energy)
            if e.type == el.RFMULT and e.sigphi != 0:
                # an RFMult read from a file drops its noise sigmas
                # (accel.cpp:1065-1071); SigPhi is settable (accel.cpp:1085)
                self.lib.ref_setfield(h, i, b"SigPhi", e.sigphi)
            if e.type == el.IDEALWIG and e.energy != el.DEFAULT_E0:
                raise ValueError("the reference's IdealWig energy is not settable from the lattice file")
        rads = {e.rad for e in elems if e.type in (el.BENDMPOLE, el.SYNCRAD, el.IDEALWIG)}
        if len(rads) > 1:
            raise ValueError("reference sets one radiation flag for the whole line")
        line.set_rad(rads.pop() if rads else 0)
        crabs = [e.on for e in elems if e.type == el.CRABCAV]
        if crabs:
            if len(set(crabs)) > 1:
                raise ValueError("reference sets one crab-cavity flag for the whole line")
            line.lib.ref_set_crab(h, int(crabs[0]))
        wakes = [e.on for e in elems if e.type == el.TRANSWAKE]
        if wakes:
            if len(set(wakes)) > 1:
                raise ValueError("reference sets one wake flag for the whole line")
            line.lib.ref_set_wake(h, int(wakes[0]))
        cav = [e.on for e in elems if e.type == el.CAVITY]
        if cav:
            if len(set(cav)) > 1:
                raise ValueError("reference sets one cavity flag for the whole line")
            line.set_cavity(cav[0])
        return line


class RefLine:
    def __init__(self, lib, h):
        self.lib, self.h = lib, h

    def __del__(self):
        try:
            self.lib.ref_free(self.h)
        except Exception:
            pass

    def set_rad(self, flag):
        self.lib.ref_set_rad(self.h, flag)

    def set_cavity(self, on):
        self.lib.ref_set_cavity(self.h, int(on))

    def set_threads(self, n):
        self.lib.ref_set_threads(n)

    def track(self, r, nturn=1):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        self.lib.ref_pass(self.h, _dp(r), r.size // 6, nturn)
        return r.reshape(-1, 6)

    def track_crn(self, r, nturn, seed=0, turn0=0, pid0=0):
        """ref_pass_crn: the reference's tracking with the oracle's normals
        in its quantum-excitation kicks (common random numbers)."""
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        if self.lib.ref_pass_crn(self.h, _dp(r), r.size // 6, nturn, seed, turn0, pid0) != 0:
            raise ValueError("ref_pass_crn failed")
        return r.reshape(-1, 6)

    def track_range(self, r, st, ed):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        self.lib.ref_pass_range(self.h, _dp(r), r.size // 6, st, ed)
        return r.reshape(-1, 6)

    def elem_pass(self, idx, r):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        self.lib.ref_elem_pass(self.h, idx, _dp(r), r.size // 6)
        return r.reshape(-1, 6)

    def setrfphase(self):
        return self.lib.ref_setrfphase(self.h)

    def getm66(self, dpp=0.0, delta=1e-8):
        m = np.zeros(36)
        self.lib.ref_getm66(self.h, dpp, delta, _dp(m))
        return m.reshape(6, 6)

    def tune(self):
        x, y = c_double(), c_double()
        self.lib.ref_calctune(self.h, ctypes.byref(x), ctypes.byref(y))
        return x.value, y.value

    def chrom(self):
        x, y = c_double(), c_double()
        self.lib.ref_calcchrom(self.h, ctypes.byref(x), ctypes.byref(y))
        return x.value, y.value

    def resp(self, kind, etype, fam1, fam2):
        out = np.zeros(4)
        self.lib.ref_calcresp(self.h, kind, etype.encode(), fam1.encode(), fam2.encode(), _dp(out))
        return out.reshape(2, 2)

    def correct(self, kind, etype, fam1, fam2, tx, ty):
        x, y = c_double(tx), c_double(ty)
        rc = self.lib.ref_correct(self.h, kind, etype.encode(), fam1.encode(), fam2.encode(),
                                  ctypes.byref(x), ctypes.byref(y))
        if rc:
            raise RuntimeError("families not found")
        return x.value, y.value

    def trackda(self, xmax, xmin, ymax, nturn, nray, nstep, negweight=False):
        rx = np.zeros(51)
        ry = np.zeros(51)
        s = self.lib.ref_trackda(self.h, xmax, xmin, ymax, nturn, nray, nstep, _dp(rx), _dp(ry), int(negweight))
        nr = min(max(nray, 3), 51)
        return s, rx[:nr], ry[:nr]

    def getco4(self, dp=0.0):
        rc = np.zeros(4)
        self.lib.ref_getco4(self.h, dp, _dp(rc))
        return rc

    def getco6(self):
        rc = np.zeros(6)
        self.lib.ref_getco6(self.h, _dp(rc))
        return rc

    def calctwiss0(self, dp=0.0):
        tw = np.zeros(7)
        self.lib.ref_calctwiss0(self.h, dp, _dp(tw))
        return tw

    def calctwiss(self, idx, dp=0.0):
        ix = np.ascontiguousarray(idx, dtype=np.int32)
        tw = np.zeros((len(ix), 7))
        self.lib.ref_calctwiss(self.h, _ip(ix), len(ix), dp, _dp(tw))
        return tw

    def calcchrom2(self):
        c = np.zeros(4)
        self.lib.ref_calcchrom2(self.h, _dp(c))
        return c

    def trackma(self, dppmax, dppmin, nturn, nstep, locs):
        lc = np.ascontiguousarray(locs, dtype=np.int32)
        p, m = np.zeros(len(lc)), np.zeros(len(lc))
        avg = self.lib.ref_trackma(self.h, dppmax, dppmin, nturn, nstep, _ip(lc), len(lc), _dp(p), _dp(m))
        return avg, p, m

    def tracktune(self, nturn, x0, y0, xlow=0.0, xhi=0.5, ylow=0.0, yhi=0.5):
        t = np.zeros(2)
        self.lib.ref_tracktune(self.h, nturn, x0, y0, xlow, xhi, ylow, yhi, _dp(t))
        return t

    def trackdiffusion(self, nturn, x0, y0, xlow=0.0, xhi=0.5, ylow=0.0, yhi=0.5):
        x0 = np.array(x0, dtype=np.float64)
        y0 = np.array(y0, dtype=np.float64)
        nux, nuy, dif = np.zeros(len(x0)), np.zeros(len(x0)), np.zeros(len(x0))
        tot = self.lib.ref_trackdiffusion(self.h, nturn, len(x0), _dp(x0), _dp(y0), _dp(nux), _dp(nuy), _dp(dif),
                                          xlow, xhi, ylow, yhi)
        return tot, nux, nuy, dif

    def trackdnudexy(self, nturn, xm, ym):
        d = np.zeros(4)
        self.lib.ref_trackdnudexy(self.h, nturn, xm, ym, _dp(d))
        return d

    def getfield(self, idx, name):
        return self.lib.ref_getfield(self.h, idx, name.encode())

    def get_mpole(self, idx, name, index):
        return self.lib.ref_get_mpole(self.h, idx, name.encode(), index)

    def setfield(self, idx, name, v):
        self.lib.ref_setfield(self.h, idx, name.encode(), v)


def beam_stat_ref(ref, r):
    r = np.ascontiguousarray(r, dtype=np.float64).copy()
    ar, sg, sr = np.zeros(6), np.zeros(6), np.zeros(6)
    nl = ref.lib.ref_beam_stat(_dp(r), r.size // 6, _dp(ar), _dp(sg), _dp(sr))
    return nl, ar, sg, sr

"""ctypes binding of libaccp.so (include/accp.h).

The library must exist (built by `make -C accel-plus_amd`); there is no
fallback: a missing library or a failing call raises AccpError.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

from . import elements as el

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libaccp.so")

STAGE_AUTO, STAGE_SCALAR, STAGE_LDS, STAGE_SPECIALIZED = 0, 1, 2, 3
STAGE_NAMES = {0: "auto", 1: "scalar", 2: "lds", 3: "specialized"}
STAT_LEN = 20


class AccpError(RuntimeError):
    pass


class Tuning(ctypes.Structure):
    _fields_ = [("block", c_int32), ("ppl", c_int32), ("staging", c_int32), ("compact", c_int32)]

    def __repr__(self):
        st = STAGE_NAMES.get(self.staging, self.staging)
        return f"Tuning(block={self.block}, ppl={self.ppl}, staging={st}, compact={self.compact})"

    def as_dict(self):
        return {"block": self.block, "ppl": self.ppl,
                "staging": STAGE_NAMES.get(self.staging, self.staging),
                "compact": self.compact}


class TrackOpts(ctypes.Structure):
    _fields_ = [
        ("rng_seed", c_uint64), ("turn0", c_uint64), ("pid0", c_uint64),
        ("mon_interval", c_int32), ("mon_max", c_int32),
        ("mon_out", POINTER(c_double)), ("mon_count", POINTER(c_int32)),
        ("loss_turn", POINTER(c_int32)),
        ("tuning", Tuning),
    ]


class Param(ctypes.Structure):
    _fields_ = [("field", c_int32), ("index", c_int32), ("nelem", c_int32), ("elems", POINTER(c_int32))]


# Every entry point include/accp.h declares.
EXPORTS = (
    "accp_last_error", "accp_abi_version", "accp_device_count",
    "accp_lattice_build", "accp_lattice_update", "accp_lattice_nelem", "accp_lattice_free",
    "accp_lattice_set_tuning",
    "accp_track", "accp_track_device", "accp_track_device_noalloc", "accp_beam_stat",
    "accp_batch_create", "accp_track_batched", "accp_batch_capture", "accp_batch_replay",
    "accp_autotune", "accp_batch_free",
    "accp_adjoint_create", "accp_adjoint_forward", "accp_adjoint_backward", "accp_adjoint_set_reverse",
    "accp_adjoint_free",
    "accp_lattice_get", "accp_getm66", "accp_calctune", "accp_calcchrom", "accp_calcresp", "accp_correct",
    "accp_setrfphase", "accp_trackda", "accp_getco4", "accp_getco6", "accp_calctwiss", "accp_calcchrom2",
    "accp_trackma", "accp_naff", "accp_tracktune", "accp_trackdiffusion", "accp_trackdnudexy",
)

DP = POINTER(c_double)


def _dp(a):
    return a.ctypes.data_as(DP)


class Library:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise AccpError(f"{path} is not built: run `make -C accel-plus_amd` (no CPU fallback exists)")
        L = ctypes.CDLL(path)
        self.path = path
        E = POINTER(el.Elem)
        VP = c_void_p
        L.accp_last_error.restype = c_char_p
        L.accp_abi_version.restype = c_int
        L.accp_device_count.argtypes = [POINTER(c_int)]
        L.accp_lattice_build.argtypes = [E, c_int, POINTER(VP)]
        L.accp_lattice_update.argtypes = [VP, E, c_int]
        L.accp_lattice_nelem.argtypes = [VP, POINTER(c_int)]
        L.accp_lattice_free.argtypes = [VP]
        L.accp_lattice_set_tuning.argtypes = [VP, POINTER(Tuning)]
        L.accp_track.argtypes = [VP, DP, c_int, c_int, c_int, c_int, POINTER(TrackOpts)]
        L.accp_track_device.argtypes = [VP, c_void_p, c_int, c_int, c_int, c_int, POINTER(TrackOpts), c_void_p]
        L.accp_track_device_noalloc.argtypes = [VP, c_void_p, c_int, c_int, c_int, c_int, POINTER(TrackOpts),
                                                c_void_p]
        L.accp_beam_stat.argtypes = [DP, c_int, DP]
        L.accp_batch_create.argtypes = [POINTER(VP), c_int, POINTER(VP)]
        L.accp_track_batched.argtypes = [VP, c_void_p, POINTER(c_int64), c_int, POINTER(Tuning), c_void_p, c_void_p]
        L.accp_batch_capture.argtypes = [VP, c_void_p, POINTER(c_int64), c_int, POINTER(Tuning), c_void_p]
        L.accp_batch_replay.argtypes = [VP, c_void_p]
        L.accp_autotune.argtypes = [VP, c_void_p, POINTER(c_int64), c_int, POINTER(Tuning), DP]
        L.accp_batch_free.argtypes = [VP]
        L.accp_adjoint_create.argtypes = [VP, c_int, POINTER(Param), POINTER(VP)]
        L.accp_adjoint_forward.argtypes = [VP, DP, c_int, c_int, c_int, DP]
        L.accp_adjoint_backward.argtypes = [VP, DP, DP, DP]
        L.accp_adjoint_free.argtypes = [VP]
        for nm in EXPORTS:
            if nm not in ("accp_last_error", "accp_abi_version"):
                getattr(L, nm).restype = c_int
        self.lib = L

    def check(self, rc, what=""):
        if rc != 0:
            raise AccpError(f"{what} failed ({rc}): {self.lib.accp_last_error().decode()}")

    def naff(self, signals, xlow=0.0, xhi=0.5):
        """naff of each row of `signals` (rows of a power-of-two length)."""
        d = np.ascontiguousarray(np.atleast_2d(signals), dtype=np.float64)
        out = np.zeros(d.shape[0])
        self.check(self.lib.accp_naff(_dp(d), d.shape[0], d.shape[1], c_double(xlow), c_double(xhi), _dp(out)),
                   "accp_naff")
        return out

    def device_count(self):
        c = c_int(0)
        self.check(self.lib.accp_device_count(byref(c)), "accp_device_count")
        return c.value

    def lattice(self, elems):
        return Lattice(self, elems)

    def beam_stat(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(STAT_LEN)
        self.check(self.lib.accp_beam_stat(_dp(r), r.size // 6, _dp(out)), "accp_beam_stat")
        return out


class Lattice:
    def __init__(self, lib, elems):
        self.L = lib
        self.arr = elems if isinstance(elems, ctypes.Array) else el.to_array(elems)
        h = c_void_p()
        lib.check(lib.lib.accp_lattice_build(self.arr, len(self.arr), byref(h)), "accp_lattice_build")
        self.h = h

    def set_tuning(self, tuning=None):
        """accp_lattice_set_tuning: the launch shape of calls on this lattice
        that name none (None: the library default)."""
        self.L.check(self.L.lib.accp_lattice_set_tuning(self.h, byref(tuning) if tuning is not None else None),
                     "accp_lattice_set_tuning")

    def update(self, elems):
        self.arr = elems if isinstance(elems, ctypes.Array) else el.to_array(elems)
        self.L.check(self.L.lib.accp_lattice_update(self.h, self.arr, len(self.arr)), "accp_lattice_update")

    def free(self):
        if self.h:
            self.L.lib.accp_lattice_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ---- ALine's lattice services (accp_getm66 ... accp_trackda) ----------
    TUNE, CHROM = 1, 2
    RESP_FD, RESP_ADJOINT = 0, 1

    def get(self):
        """The lattice's element records (accp_lattice_get)."""
        out = (el.Elem * len(self.arr))()
        self.L.check(self.L.lib.accp_lattice_get(self.h, out, len(out)), "accp_lattice_get")
        return out

    def getm66(self, st=0, ed=-1, dp=0.0, delta=1e-8):
        m = np.zeros(36)
        self.L.check(self.L.lib.accp_getm66(self.h, st, ed, c_double(dp), c_double(delta), _dp(m)), "accp_getm66")
        return m.reshape(6, 6)

    def calctune(self, dp=0.0, delta=1e-8):
        t = np.zeros(2)
        self.L.check(self.L.lib.accp_calctune(self.h, c_double(dp), c_double(delta), _dp(t)), "accp_calctune")
        return t

    def calcchrom(self, dp=0.0, delta=1e-8):
        c = np.zeros(2)
        self.L.check(self.L.lib.accp_calcchrom(self.h, c_double(dp), c_double(delta), _dp(c)), "accp_calcchrom")
        return c

    @staticmethod
    def _fam(f):
        a = np.ascontiguousarray(f, dtype=np.int32)
        return a, a.ctypes.data_as(POINTER(c_int32)), len(a)

    def calcresp(self, quantity, fam1, fam2, method=0):
        """resp[k, p] = d(tune or chromaticity of plane p)/d(knob k)."""
        a1, p1, n1 = self._fam(fam1)
        a2, p2, n2 = self._fam(fam2)
        r = np.zeros(4)
        self.L.check(self.L.lib.accp_calcresp(self.h, quantity, method, p1, n1, p2, n2, _dp(r)), "accp_calcresp")
        return r.reshape(2, 2)

    def correct(self, quantity, fam1, fam2, target, method=0):
        """One correction step; returns the tunes (chromaticities) reached."""
        a1, p1, n1 = self._fam(fam1)
        a2, p2, n2 = self._fam(fam2)
        t = np.array(target, dtype=np.float64)
        self.L.check(self.L.lib.accp_correct(self.h, quantity, method, p1, n1, p2, n2, _dp(t)), "accp_correct")
        self.arr = self.get()
        return t

    def setrfphase(self):
        ph = c_double(0.0)
        self.L.check(self.L.lib.accp_setrfphase(self.h, byref(ph)), "accp_setrfphase")
        self.arr = self.get()
        return ph.value

    def trackda(self, xmax, xmin, ymax, nturn, nray, nstep, negweight=False, loss_turn=False):
        """Returns (area, rx[nray], ry[nray]) and, if loss_turn, the loss turn
        of every grid point."""
        rx, ry = np.zeros(51), np.zeros(51)
        area = c_double(0.0)
        nr = min(max(nray, 3), 51)
        lt = np.zeros(nr * max(nstep, 2), dtype=np.int32) if loss_turn else None
        self.L.check(self.L.lib.accp_trackda(
            self.h, c_double(xmax), c_double(xmin), c_double(ymax), nturn, nray, nstep, int(negweight),
            _dp(rx), _dp(ry), byref(area), lt.ctypes.data_as(POINTER(c_int32)) if lt is not None else None),
            "accp_trackda")
        out = (area.value, rx[:nr], ry[:nr])
        return out + (lt,) if loss_turn else out

    def getco4(self, dp=0.0):
        rc = np.zeros(4)
        self.L.check(self.L.lib.accp_getco4(self.h, c_double(dp), _dp(rc)), "accp_getco4")
        self.arr = self.get()
        return rc

    def getco6(self):
        rc = np.zeros(6)
        self.L.check(self.L.lib.accp_getco6(self.h, _dp(rc)), "accp_getco6")
        self.arr = self.get()
        return rc

    def calctwiss(self, idx=(), dp=0.0, delta=1e-8):
        """Rows {s, psix, psiy, betx, alfx, bety, alfy}: one at the ring start
        when idx is empty, else one per listed element."""
        ix = np.ascontiguousarray(idx, dtype=np.int32)
        tw = np.zeros((max(len(ix), 1), 7))
        self.L.check(self.L.lib.accp_calctwiss(self.h, ix.ctypes.data_as(POINTER(c_int32)), len(ix), c_double(dp),
                                               c_double(delta), _dp(tw)), "accp_calctwiss")
        return tw[0] if len(ix) == 0 else tw

    def calcchrom2(self):
        c = np.zeros(4)
        self.L.check(self.L.lib.accp_calcchrom2(self.h, _dp(c)), "accp_calcchrom2")
        return c

    def trackma(self, dppmax, dppmin, nturn, nstep, locs):
        lc = np.ascontiguousarray(locs, dtype=np.int32)
        p, m, avg = np.zeros(len(lc)), np.zeros(len(lc)), c_double(0.0)
        self.L.check(self.L.lib.accp_trackma(self.h, c_double(dppmax), c_double(dppmin), nturn, nstep,
                                             lc.ctypes.data_as(POINTER(c_int32)), len(lc), _dp(p), _dp(m),
                                             byref(avg)), "accp_trackma")
        return avg.value, p, m

    def tracktune(self, nturn, x0, y0, xlow=0.0, xhi=0.5, ylow=0.0, yhi=0.5):
        t = np.zeros(2)
        self.L.check(self.L.lib.accp_tracktune(self.h, nturn, c_double(x0), c_double(y0), c_double(xlow),
                                               c_double(xhi), c_double(ylow), c_double(yhi), _dp(t)),
                     "accp_tracktune")
        return t

    def trackdiffusion(self, nturn, x0, y0, xlow=0.0, xhi=0.5, ylow=0.0, yhi=0.5):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        nux, nuy, dif, tot = np.zeros(len(x0)), np.zeros(len(x0)), np.zeros(len(x0)), c_double(0.0)
        self.L.check(self.L.lib.accp_trackdiffusion(self.h, nturn, len(x0), _dp(x0), _dp(y0), _dp(nux), _dp(nuy),
                                                    _dp(dif), c_double(xlow), c_double(xhi), c_double(ylow),
                                                    c_double(yhi), byref(tot)), "accp_trackdiffusion")
        return tot.value, nux, nuy, dif

    def trackdnudexy(self, nturn, xm, ym, rows=False):
        d, rw = np.zeros(4), np.zeros(36)
        self.L.check(self.L.lib.accp_trackdnudexy(self.h, nturn, c_double(xm), c_double(ym), _dp(d), _dp(rw)),
                     "accp_trackdnudexy")
        return (d, rw.reshape(12, 3)) if rows else d

    def track(self, r, nturn=1, st=0, ed=-1, seed=0, turn0=0, pid0=0, tuning=None,
              mon_interval=0, mon_max=0, loss_turn=False):
        """ALine::pass semantics on a host AoS array.  Returns r_out, plus the
        monitor records if mon_interval > 0, plus the per-particle loss turns
        if loss_turn."""
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        o = TrackOpts()
        o.rng_seed, o.turn0, o.pid0 = seed, turn0, pid0
        if tuning is not None:
            o.tuning = tuning
        lt = None
        if loss_turn:
            lt = np.zeros(r.size // 6, dtype=np.int32)
            o.loss_turn = lt.ctypes.data_as(POINTER(c_int32))
        recs = None
        cnt = c_int32(0)
        if mon_interval > 0:
            mon_max = mon_max or (nturn + mon_interval - 1) // mon_interval
            recs = np.zeros((mon_max, STAT_LEN))
            o.mon_interval, o.mon_max = mon_interval, mon_max
            o.mon_out = _dp(recs)
            o.mon_count = ctypes.pointer(cnt)
        self.L.check(self.L.lib.accp_track(self.h, _dp(r), r.size // 6, nturn, st, ed, byref(o)), "accp_track")
        out = (r.reshape(-1, 6),)
        if recs is not None:
            out += (recs[: cnt.value],)
        if lt is not None:
            out += (lt,)
        return out if len(out) > 1 else out[0]

    def track_device(self, d_ptr, n, nturn, st=0, ed=-1, seed=0, turn0=0, pid0=0, tuning=None, stream=None,
                     d_loss_turn=None, noalloc=False):
        """accp_track_device (noalloc: accp_track_device_noalloc)."""
        o = TrackOpts()
        o.rng_seed, o.turn0, o.pid0 = seed, turn0, pid0
        if d_loss_turn:
            o.loss_turn = ctypes.cast(c_void_p(d_loss_turn), POINTER(c_int32))
        if tuning is not None:
            o.tuning = tuning
        fn = self.L.lib.accp_track_device_noalloc if noalloc else self.L.lib.accp_track_device
        self.L.check(fn(self.h, c_void_p(d_ptr), n, nturn, st, ed, byref(o), c_void_p(stream) if stream else None),
                     "accp_track_device_noalloc" if noalloc else "accp_track_device")


class Batch:
    def __init__(self, lib, lattices):
        self.L = lib
        self.lats = lattices
        hs = (c_void_p * len(lattices))(*[x.h.value for x in lattices])
        h = c_void_p()
        lib.check(lib.lib.accp_batch_create(hs, len(lattices), byref(h)), "accp_batch_create")
        self.h = h

    def track(self, d_ptr, offsets, nturn, tuning=None, stream=None, d_loss_turn=None):
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        self._offs = offs
        t = byref(tuning) if tuning is not None else None
        self.L.check(self.L.lib.accp_track_batched(self.h, c_void_p(d_ptr), offs.ctypes.data_as(POINTER(c_int64)),
                                                   nturn, t, c_void_p(d_loss_turn) if d_loss_turn else None,
                                                   c_void_p(stream) if stream else None),
                     "accp_track_batched")

    def capture(self, d_ptr, offsets, nturn, tuning=None, d_loss_turn=None):
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        self._offs = offs
        t = byref(tuning) if tuning is not None else None
        self.L.check(self.L.lib.accp_batch_capture(self.h, c_void_p(d_ptr), offs.ctypes.data_as(POINTER(c_int64)),
                                                   nturn, t, c_void_p(d_loss_turn) if d_loss_turn else None),
                     "accp_batch_capture")

    def replay(self, stream=None):
        self.L.check(self.L.lib.accp_batch_replay(self.h, c_void_p(stream) if stream else None), "accp_batch_replay")

    def autotune(self, d_ptr, offsets, nturn_probe):
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        best = Tuning()
        ms = c_double()
        self.L.check(self.L.lib.accp_autotune(self.h, c_void_p(d_ptr), offs.ctypes.data_as(POINTER(c_int64)),
                                              nturn_probe, byref(best), byref(ms)), "accp_autotune")
        return best, ms.value

    def free(self):
        if self.h:
            self.L.lib.accp_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_LIB = None


def load(path=None):
    global _LIB
    if path is not None:
        return Library(path)
    if _LIB is None:
        _LIB = Library()
    return _LIB


class Adjoint:
    """accp_adjoint_* : gradient of a loss on the final coordinates with
    respect to family knobs.  params: list of (field, index, element list),
    field one of elements.FIELD_K / FIELD_POLYB / FIELD_POLYA."""

    def __init__(self, lib, lattice, params):
        self.L = lib
        self.lattice = lattice
        self._members = [np.ascontiguousarray(m, dtype=np.int32) for _, _, m in params]
        arr = (Param * max(len(params), 1))()
        for k, (field, index, _) in enumerate(params):
            arr[k] = Param(field, index, len(self._members[k]),
                           self._members[k].ctypes.data_as(POINTER(c_int32)))
        h = c_void_p()
        lib.check(lib.lib.accp_adjoint_create(lattice.h, len(params), arr, byref(h)), "accp_adjoint_create")
        self.h = h
        self.nparam = len(params)

    REVERSE_AUTO, REVERSE_SEQUENTIAL, REVERSE_SEGMENT_JACOBIAN = 0, 1, 2

    def set_reverse(self, mode):
        """accp_adjoint_set_reverse: AUTO, SEQUENTIAL or SEGMENT_JACOBIAN."""
        self.L.check(self.L.lib.accp_adjoint_set_reverse(self.h, mode), "accp_adjoint_set_reverse")

    def forward(self, r0, nturn, ckpt_every=16):
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        out = np.empty_like(r0)
        self.L.check(self.L.lib.accp_adjoint_forward(self.h, _dp(r0), r0.size // 6, nturn, ckpt_every, _dp(out)),
                     "accp_adjoint_forward")
        return out.reshape(-1, 6)

    def backward(self, dl_dr_out):
        g = np.ascontiguousarray(dl_dr_out, dtype=np.float64)
        grad = np.zeros(max(self.nparam, 1))
        dr0 = np.empty_like(g)
        self.L.check(self.L.lib.accp_adjoint_backward(self.h, _dp(g), _dp(grad), _dp(dr0)), "accp_adjoint_backward")
        return grad[: self.nparam], dr0.reshape(-1, 6)

    def free(self):
        if self.h:
            self.L.lib.accp_adjoint_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

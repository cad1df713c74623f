"""Element records of the accp C ABI and the reference lattice-file grammar.

`Elem` mirrors `accp_elem` in include/accp.h field for field.  The writer
emits the Accel+ lattice-file format that AccelElement::deserialize and the
per-type create() parsers read (src/lib/accel/accel.cpp:67-133, :147-157,
:184-192, :236-244, :299-307, :394-407, :481-489, :560-573, :718-736,
:860-869), so the compiled reference and this backend can load the same file.
"""
import ctypes
from ctypes import c_char, c_double, c_int32

import numpy as np

MAX_MPOLE = 10

(DRIFT, QUAD, CORRECTOR, APERTURE, MONITOR, BEND, BENDMPOLE, STRMPOLE, CAVITY, LONGWAKE, MATRIX, SYNCRAD, CRABCAV,
 RFMULT, IDEALWIG, TRANSWAKE, WIGTABLE) = range(1, 18)
TYPE_NAMES = {
    DRIFT: "Drift", QUAD: "Quad", CORRECTOR: "Corrector", APERTURE: "Aperture",
    MONITOR: "Monitor", BEND: "Bend", BENDMPOLE: "BendMPole", STRMPOLE: "StrMPole",
    CAVITY: "Cavity", LONGWAKE: "LongWake", MATRIX: "Matrix", SYNCRAD: "SyncRad", CRABCAV: "CrabCav",
    RFMULT: "RFMult", IDEALWIG: "IdealWig", TRANSWAKE: "TransWake", WIGTABLE: "WigTable",
}
WIGTAB_MAX_X, WIGTAB_MAX_Y = 101, 21  # accel.h:401-402
RAD_OFF, RAD_DAMP, RAD_QE, RAD_LUMP, RAD_LUMP_DAMP, RAD_LUMP_QE = 0, 1, 2, 3, 4, 5
FIELD_K, FIELD_POLYB, FIELD_POLYA = 1, 2, 3
DEFAULT_E0 = 3000.0  # MeV, accel.h:30
C0 = 2.99792458e8  # accel.h:29


class Elem(ctypes.Structure):
    _fields_ = [
        ("type", c_int32), ("nslice", c_int32), ("max_npole", c_int32),
        ("rad", c_int32), ("on", c_int32), ("nbins", c_int32),
        ("len", c_double), ("k", c_double), ("theta", c_double),
        ("e1", c_double), ("e2", c_double), ("energy", c_double),
        ("volt", c_double), ("freq", c_double), ("phis", c_double),
        ("kickx", c_double), ("kicky", c_double),
        ("xlim", c_double * 2), ("ylim", c_double * 2),
        ("polyB", c_double * MAX_MPOLE), ("polyA", c_double * MAX_MPOLE),
        ("wake_rs", c_double), ("wake_q", c_double), ("wake_fr", c_double),
        ("wake_charge", c_double), ("wake_zmin", c_double), ("wake_zmax", c_double),
        ("mat", c_double * 36),
        ("u0", c_double), ("jx", c_double), ("jy", c_double),
        ("emitx", c_double), ("emity", c_double), ("sigdpp", c_double), ("betax", c_double), ("betay", c_double),
        ("devoltx", c_double), ("devolty", c_double), ("sigphi", c_double), ("sigdvv", c_double),
        ("wig_lambda", c_double), ("wig_h", c_double),
        ("tw_w", c_double * 2), ("tw_fr", c_double * 2), ("tw_q", c_double * 2),
        ("wig_type", c_int32), ("wt_mx", c_int32), ("wt_my", c_int32), ("reserved", c_int32),
        ("wt_table", ctypes.POINTER(c_double)),
        ("famname", c_char * 32),
    ]

    @property
    def name(self):
        return self.famname.decode()

    def copy(self):
        e = Elem()
        ctypes.pointer(e)[0] = self
        return e

    def __repr__(self):
        return f"Elem({TYPE_NAMES.get(self.type, self.type)} {self.name} L={self.len})"


def _mk(t, name, L, **kw):
    e = Elem()
    e.type = t
    e.len = float(L)
    e.famname = name.encode()[:31]
    for k, v in kw.items():
        if k in ("polyB", "polyA"):
            arr = getattr(e, k)
            for i, x in enumerate(v):
                arr[i] = float(x)
        elif k in ("xlim", "ylim", "tw_w", "tw_fr", "tw_q"):
            arr = getattr(e, k)
            arr[0], arr[1] = float(v[0]), float(v[1])
        elif k == "mat":
            arr = getattr(e, k)
            for i, x in enumerate(v):
                arr[i] = float(x)
        else:
            setattr(e, k, v)
    return e


def drift(L, name="D"):
    return _mk(DRIFT, name, L)


def quad(L, K, name="Q"):
    return _mk(QUAD, name, L, k=float(K))


def corrector(L, kx, ky, name="COR"):
    return _mk(CORRECTOR, name, L, kickx=float(kx), kicky=float(ky))


def aperture(L, xlim, ylim, name="AP"):
    return _mk(APERTURE, name, L, xlim=xlim, ylim=ylim)


def monitor(name="MON"):
    return _mk(MONITOR, name, 0.0)


def bend(L, theta, e1=0.0, e2=0.0, K=0.0, name="B"):
    return _mk(BEND, name, L, theta=float(theta), e1=float(e1), e2=float(e2), k=float(K))


def bendmpole(L, theta, e1=0.0, e2=0.0, polyB=(0.0,), name="BM", nslice=10, energy=DEFAULT_E0, rad=0):
    return _mk(BENDMPOLE, name, L, theta=float(theta), e1=float(e1), e2=float(e2),
               polyB=polyB, max_npole=len(polyB), nslice=nslice, energy=float(energy), rad=rad)


def strmpole(L, polyB, polyA=None, name="SM", nslice=10, max_npole=None):
    n = max_npole if max_npole is not None else len(polyB)
    polyA = polyA if polyA is not None else [0.0] * n
    return _mk(STRMPOLE, name, L, polyB=polyB, polyA=polyA, max_npole=n, nslice=nslice)


def cavity(L, volt, freq, phis=0.0, energy=DEFAULT_E0, on=1, name="RF"):
    return _mk(CAVITY, name, L, volt=float(volt), freq=float(freq), phis=float(phis),
               energy=float(energy), on=on)


def matrix(L, m, name="MAT"):
    """AMatrix: a 6x6 map (row-major, 36 numbers) between two half drifts."""
    return _mk(MATRIX, name, L, mat=[float(x) for x in m])


def syncrad(L, energy, u0, jx, jy, emitx, emity, sigdpp, betax, betay, rad=RAD_OFF, name="SR"):
    """ASyncRad: lumped radiation damping and excitation."""
    return _mk(SYNCRAD, name, L, energy=float(energy), u0=float(u0), jx=float(jx), jy=float(jy),
               emitx=float(emitx), emity=float(emity), sigdpp=float(sigdpp), betax=float(betax),
               betay=float(betay), rad=rad)


def crabcav(L, devoltx, devolty, freq, phis, energy=DEFAULT_E0, sigphi=0.0, sigdvv=0.0, on=1, name="CRAB"):
    """ACrabCav: transverse deflecting cavity with phase and amplitude noise."""
    return _mk(CRABCAV, name, L, devoltx=float(devoltx), devolty=float(devolty), freq=float(freq),
               phis=float(phis), energy=float(energy), sigphi=float(sigphi), sigdvv=float(sigdvv), on=on)


def rfmult(L, freq, phis, polyB, polyA=None, energy=DEFAULT_E0, sigphi=0.0, sigdvv=0.0, name="RFM"):
    """ARFMult: RF multipole, fields B_n, A_n modulated by cos(-k z + phis)."""
    n = len(polyB)
    return _mk(RFMULT, name, L, freq=float(freq), phis=float(phis), energy=float(energy), sigphi=float(sigphi),
               sigdvv=float(sigdvv), polyB=polyB, polyA=polyA if polyA is not None else [0.0] * n, max_npole=n)


def idealwig(L, nslice, lam, h, wig_type=0, energy=DEFAULT_E0, rad=RAD_OFF, name="WIG"):
    """AIdealWig: wiggler of period lam and peak curvature h, horizontal (0) or vertical (1)."""
    return _mk(IDEALWIG, name, L, nslice=int(nslice), wig_lambda=float(lam), wig_h=float(h), wig_type=int(wig_type),
               energy=float(energy), rad=rad)


def transwake(L, ne, wx, frx, qx, wy, fry, qy, energy=DEFAULT_E0, on=1, name="TWAKE"):
    """ATransWake: transverse resonator wake of the whole beam; ne = bunch
    charge [nC], W per plane, resonant frequency [GHz], quality factor."""
    return _mk(TRANSWAKE, name, L, wake_charge=float(ne), tw_w=(wx, wy), tw_fr=(frx, fry), tw_q=(qx, qy),
               energy=float(energy), on=on)


def longwake(rs, q, fr, charge, nbins, zmin, zmax, energy=DEFAULT_E0, on=1, name="WAKE"):
    return _mk(LONGWAKE, name, 0.0, wake_rs=float(rs), wake_q=float(q), wake_fr=float(fr),
               wake_charge=float(charge), nbins=int(nbins), wake_zmin=float(zmin),
               wake_zmax=float(zmax), energy=float(energy), on=on)


def to_array(elems):
    arr = (Elem * len(elems))()
    for i, e in enumerate(elems):
        arr[i] = e
    return arr


def copy_list(elems):
    return [e.copy() for e in elems]


def _g(v):
    return "%.17g" % v


# WigTable buffers stay alive as long as the process: records are copied by
# value into ctypes arrays, which do not keep Python references.
_TABLES = []


def wigtable(L, nslice, xl, yl, xkick, ykick, name="WT"):
    """AWigTable: kicks [rad] tabulated on the grid xl (mx points) by yl (my
    points), xkick[iy][ix] and ykick[iy][ix], interpolated bilinearly;
    nslice drift-kick-drift slices for L > 0, one full kick for L = 0."""
    xl, yl = np.asarray(xl, float), np.asarray(yl, float)
    xk, yk = np.asarray(xkick, float), np.asarray(ykick, float)
    mx, my = len(xl), len(yl)
    if not (2 <= mx <= WIGTAB_MAX_X and 2 <= my <= WIGTAB_MAX_Y) or xk.shape != (my, mx) or yk.shape != (my, mx):
        raise ValueError("WigTable: need 2<=mx<=101, 2<=my<=21 and (my, mx) kick tables")
    buf = np.ascontiguousarray(np.concatenate([xl, yl, xk.ravel(), yk.ravel()]))
    _TABLES.append(buf)
    return _mk(WIGTABLE, name, L, nslice=int(nslice), wt_mx=mx, wt_my=my,
               wt_table=buf.ctypes.data_as(ctypes.POINTER(c_double)))


def wigtable_arrays(e):
    """(xl, yl, xkick, ykick) of a WigTable record, as copies."""
    mx, my = e.wt_mx, e.wt_my
    v = np.ctypeslib.as_array(e.wt_table, shape=(mx + my + 2 * mx * my,)).copy()
    return v[:mx], v[mx:mx + my], v[mx + my:mx + my + mx * my].reshape(my, mx), v[mx + my + mx * my:].reshape(my, mx)


def write_lattice(elems, path):
    """Write `elems` in the Accel+ lattice-file grammar (one record per
    element, multipole coefficient lines after StrMPole/BendMPole headers,
    terminated by RETURN).  LongWake has no reference record and is refused."""
    lines = []
    for e in elems:
        n = e.name
        t = e.type
        if t == DRIFT:
            lines.append(f"{n} : Drift {_g(e.len)}")
        elif t == QUAD:
            lines.append(f"{n} : Quad {_g(e.len)} {_g(e.k)}")
        elif t == CORRECTOR:
            lines.append(f"{n} : Corrector {_g(e.len)} {_g(e.kickx)} {_g(e.kicky)}")
        elif t == APERTURE:
            lines.append(f"{n} : Aperture {_g(e.len)} {_g(e.xlim[0])} {_g(e.xlim[1])} "
                         f"{_g(e.ylim[0])} {_g(e.ylim[1])}")
        elif t == MONITOR:
            lines.append(f"{n} : Monitor 0 1 1")
        elif t == BEND:
            lines.append(f"{n} : Bend {_g(e.len)} {_g(e.theta)} {_g(e.e1)} {_g(e.e2)} {_g(e.k)}")
        elif t == BENDMPOLE:
            lines.append(f"{n} : BendMPole {_g(e.len)} {_g(e.theta)} {_g(e.e1)} {_g(e.e2)} {e.max_npole}")
            lines.append("   " + " ".join(_g(e.polyB[i]) for i in range(e.max_npole)))
        elif t == STRMPOLE:
            lines.append(f"{n} : StrMPole {_g(e.len)} {e.max_npole}")
            lines.append("   " + " ".join(_g(e.polyB[i]) for i in range(e.max_npole)))
            lines.append("   " + " ".join(_g(e.polyA[i]) for i in range(e.max_npole)))
        elif t == CAVITY:
            lines.append(f"{n} : Cavity {_g(e.len)} {_g(e.volt)} {_g(e.freq)} {_g(e.phis)} {_g(e.energy)}")
        elif t == MATRIX:
            lines.append(f"{n} : Matrix {_g(e.len)}")
            for i in range(6):
                lines.append("   " + " ".join(_g(e.mat[6 * i + j]) for j in range(6)))
        elif t == SYNCRAD:
            lines.append(f"{n} : SyncRad {_g(e.len)} {_g(e.energy)} {_g(e.u0)} {_g(e.jx)} {_g(e.jy)}")
            lines.append(f"   {_g(e.emitx)} {_g(e.emity)} {_g(e.sigdpp)} {_g(e.betax)} {_g(e.betay)}")
        elif t == RFMULT:
            lines.append(f"{n} : RFMult {_g(e.len)} {_g(e.freq)} {_g(e.phis)} {_g(e.energy)} {_g(e.sigphi)} "
                         f"{_g(e.sigdvv)} {e.max_npole}")
            lines.append("   " + " ".join(_g(e.polyB[i]) for i in range(e.max_npole)))
            lines.append("   " + " ".join(_g(e.polyA[i]) for i in range(e.max_npole)))
        elif t == IDEALWIG:
            lines.append(f"{n} : IdealWig {_g(e.len)} {e.nslice} {_g(e.wig_lambda)} {_g(e.wig_h)} {e.wig_type}")
        elif t == TRANSWAKE:
            lines.append(f"{n} : TransWake {_g(e.len)} {_g(e.energy)} {_g(e.wake_charge)} {_g(e.tw_w[0])} "
                         f"{_g(e.tw_fr[0])} {_g(e.tw_q[0])} {_g(e.tw_w[1])} {_g(e.tw_fr[1])} {_g(e.tw_q[1])}")
        elif t == WIGTABLE:
            # AWigTable::create (accel.cpp:1298-1335) appends each line to one
            # stream, so every line ends in a blank, as serialize writes them
            xl, yl, xk, yk = wigtable_arrays(e)
            lines.append(f"{n} : WigTable {_g(e.len)} {e.nslice} {e.wt_mx} {e.wt_my}")
            lines.append("".join(_g(v) + "\t" for v in xl))
            lines.append("".join(_g(v) + "\t" for v in yl))
            lines += ["".join(_g(v) + "\t" for v in row) for row in xk]
            lines.append("")
            lines += ["".join(_g(v) + "\t" for v in row) for row in yk]
        elif t == CRABCAV:
            lines.append(f"{n} : CrabCav {_g(e.len)} {_g(e.devoltx)} {_g(e.devolty)} {_g(e.freq)} {_g(e.phis)} "
                         f"{_g(e.energy)} {_g(e.sigphi)} {_g(e.sigdvv)}")
        else:
            raise ValueError(f"element type {t} has no reference lattice record")
    lines.append("RETURN")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_plain_beam(r, path):
    """Particles as the `Distribution: Load f, format=plain` reader expects:
    six numbers per line (process.cpp:1519-1545), full precision."""
    r = np.asarray(r, dtype=np.float64).reshape(-1, 6)
    with open(path, "w") as f:
        for row in r:
            f.write(" ".join(_g(v) for v in row) + "\n")

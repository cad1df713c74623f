"""Algorithmic FP64 flop counts per particle-turn of a lattice.

Convention: + - * / each count 1; sqrt, sin, cos, sinh, cosh, tan, exp,
log are counted separately (not in these totals); comparisons and integer
work are free.  Counts follow the reference primitives
(src/lib/accel/atpass.cpp) term by term:

  fastdrift (:416-428)            12
  ADrift::pass (accel.cpp:170)    2 + 12
  quad6 (:15-72)                  57
  bend6 (:76-187)                 102, edge (:190-196) 5 each
  ATbendhxdrift6 (:343-355)       46
  bndthinkick (:199-235)          44 + 7 per Horner step
  bndthinkickrad (:255-339)       113 + 7 per Horner step (+6 with QE)
  strthinkick (:358-414)          4 + 8 per Horner step
  StrMPole slice (accel.cpp:833-842)  4 + 4*fastdrift + 3*strthinkick
  ACavity::pass (:903-923)        7 (+ 2*(2+12) with length)

mode="reference": the reference's own algorithm (StrMPole Horner always of
order 9, accel.cpp:826/:837-841; full bndthinkick; edges always applied).
mode="useful": the work the GPU kernel does (Horner from the highest
non-zero order, curvature terms dropped when PolyB[1] = PolyB[2] = 0,
zero-angle edges skipped).
"""
from . import elements as el

FASTDRIFT = 12
HXDRIFT = 46
EDGE = 5


def _hi(e, with_a=True):
    hi = -1
    for i in range(e.max_npole):
        if e.polyB[i] != 0 or (with_a and e.polyA[i] != 0):
            hi = i
    return hi


def element_flops(e, mode="useful"):
    t = e.type
    ref = mode == "reference"
    if t == el.DRIFT:
        return 2 + FASTDRIFT
    if t == el.QUAD:
        return 57 if e.k != 0 else 12
    if t == el.CORRECTOR:
        return 4 + (2 * (2 + FASTDRIFT) if e.len != 0 else 0)
    if t == el.APERTURE:
        return 2 + FASTDRIFT if e.len != 0 else 0
    if t == el.BEND:
        return 102 + 2 * EDGE
    if t == el.BENDMPOLE:
        mo = e.max_npole - 1
        steps = mo if ref else max(_hi(e, with_a=False), 0)
        plain = e.max_npole < 2 or (e.polyB[1] == 0 and (e.max_npole < 3 or e.polyB[2] == 0))
        if e.rad == el.RAD_QE:
            kick = 113 + 6 + 7 * steps
            per_slice = 2 * HXDRIFT + kick
        else:
            if e.rad == el.RAD_DAMP:
                kick = 113 + 7 * steps
            elif plain and not ref:
                kick = 9 + 7 * steps + (4 if steps > 0 else 0)
            else:
                kick = 44 + 7 * steps
            per_slice = 4 * HXDRIFT + 3 * kick
        edges = 2 * EDGE if ref else EDGE * ((e.e1 != 0) + (e.e2 != 0))
        return e.nslice * per_slice + edges
    if t == el.STRMPOLE:
        order = 9 if ref else _hi(e)
        kick = 4 + 8 * order if order >= 0 else 0
        if e.len == 0:
            return kick
        return e.nslice * (4 + 4 * FASTDRIFT + 3 * kick)
    if t == el.CAVITY:
        return (7 if e.on else 0) + (2 * (2 + FASTDRIFT) if e.len != 0 else 0)
    drifts = 2 * (2 + FASTDRIFT) if e.len != 0 else 0
    if t == el.MATRIX:  # AMatrix::pass (accel.cpp:1715-1719): 36 multiplies, 30 adds
        return 66 + drifts
    if t == el.SYNCRAD:  # ASyncRad::pass (accel.cpp:1830-1848)
        damp = e.rad in (el.RAD_LUMP, el.RAD_LUMP_DAMP)
        qe = e.rad in (el.RAD_LUMP, el.RAD_LUMP_QE)
        return (3 if damp else 0) + (6 if qe else 0) + drifts
    if t == el.CRABCAV:  # ACrabCav::pass (accel.cpp:977-995)
        return (15 if e.on else 0) + drifts
    if t == el.WIGTABLE:  # AWigTable::pass (accel.cpp:1338-1382)
        # interp2 (interp_2d.h:10-20): cell fractions 6, weights 2, corners 8, sum 3
        kick = 2 * 19
        if e.len == 0:
            return kick + 2
        return e.nslice * (2 * (3 + FASTDRIFT) + kick + 4)
    return 0


def flops_per_turn(elems, mode="useful"):
    return sum(element_flops(e, mode) for e in elems)

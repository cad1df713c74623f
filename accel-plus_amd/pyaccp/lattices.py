"""Synthetic lattices and beams of the benchmark workloads.

Ring: 20 identical 8.0 m FODO cells (C = 160 m), each
    D1.0 QF D0.5 SF D0.5 B D0.5 QD D0.5 SD B D1.0
QF/QD: Quad L=0.3, K=+1.2/-1.1; SF/SD: StrMPole L=0.2, PolyB[2]=+20/-25;
B: BendMPole L=1.5, angle 2*pi/40, MaxNPole=1, PolyB=[0] (nslice 10,
3000 MeV: the reference defaults, accel.h:195-196).

The DA lattice adds one RF cavity (1.5 MV, harmonic 300), a zero-length
aperture (|x|<25 mm, |y|<12 mm) and, after every Quad and bend, one thin
StrMPole carrying seeded relative multipole errors b3..b10, a3..a10
(sigma 1e-4 at r0 = 10 mm): the reference's AQuad/ABendMPole cannot carry
them (accel.h:89-102, :190), AStrMPole with len==0 applies one integrated
kick (accel.cpp:824-827).
"""
import math

import numpy as np

from . import elements as el

N_CELLS = 20
CIRC = 160.0
HARMONIC = 300
RF_FREQ = HARMONIC * el.C0 / CIRC
ERR_R0 = 0.01
ERR_SIGMA = 1e-4
BEAM_SIGMAS = np.array([100e-6, 10e-6, 20e-6, 5e-6, 3e-3, 1e-3])  # x, px, y, py, z, delta


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def fodo_ring(n_cells=N_CELLS, kqf=1.2, kqd=-1.1, sf=20.0, sd=-25.0, bend_type="BendMPole"):
    """The 240-element ring: a list of Elem."""
    theta = 2 * math.pi / (2 * n_cells)
    out = []
    for _ in range(n_cells):
        b = (el.bendmpole(1.5, theta, polyB=[0.0], name="B") if bend_type == "BendMPole"
             else el.bend(1.5, theta, name="B"))
        out += [
            el.drift(1.0, "D1"), el.quad(0.3, kqf, "QF"), el.drift(0.5, "D2"),
            el.strmpole(0.2, [0.0, 0.0, sf], name="SF"), el.drift(0.5, "D2"), b,
            el.drift(0.5, "D2"), el.quad(0.3, kqd, "QD"), el.drift(0.5, "D2"),
            el.strmpole(0.2, [0.0, 0.0, sd], name="SD"), b.copy(), el.drift(1.0, "D1"),
        ]
    return out


def combined_function_ring(gradient=-0.1, sext=2.0):
    """The FODO ring with combined-function bends: BendMPole with PolyB =
    [0, gradient, sext], so the curvature terms of PolyB[1] and PolyB[2]
    enter every kick (at the defaults cos mu = -0.13 and 0.50)."""
    return [el.bendmpole(1.5, e.theta, polyB=[0.0, gradient, sext], name="B") if e.type == el.BENDMPOLE else e
            for e in fodo_ring()]


def _error_kick(rng, strength, is_quad):
    b = rng.normal(0.0, ERR_SIGMA, 8)
    a = rng.normal(0.0, ERR_SIGMA, 8)
    pb = [0.0] * el.MAX_MPOLE
    pa = [0.0] * el.MAX_MPOLE
    for i, n in enumerate(range(3, 11)):
        scale = strength / (ERR_R0 ** (n - 2) if is_quad else ERR_R0 ** (n - 1))
        pb[n - 1] = scale * b[i]
        pa[n - 1] = scale * a[i]
    return el.strmpole(0.0, pb, pa, name="EMP", max_npole=el.MAX_MPOLE)


def da_ring(seed=0, extra_quads=False, strength_errors=False, rf_on=1, phis=0.0):
    """Config-1 lattice (seed 0) and its error seeds.

    strength_errors: per-element Quad K*(1+N(0,1e-3)) and SF/SD
    PolyB[2]*(1+N(0,5e-3)) drawn after the multipole errors (batched seeds).
    extra_quads: 20 Quads (L=0.1) at the cell midpoints (the first 0.1 m of
    the drift after the first bend), families QM1..QM4 round-robin with
    K = +0.02, -0.02, +0.015, -0.015 (adjoint workload)."""
    rng = _rng(seed)
    out = []
    for e in fodo_ring():
        out.append(e)
        if e.type == el.QUAD:
            out.append(_error_kick(rng, e.k * e.len, True))
        elif e.type in (el.BENDMPOLE, el.BEND):
            out.append(_error_kick(rng, e.theta, False))
    if strength_errors:
        for e in out:
            if e.type == el.QUAD:
                e.k = e.k * (1.0 + rng.normal(0.0, 1e-3))
            elif e.type == el.STRMPOLE and e.name in ("SF", "SD"):
                e.polyB[2] = e.polyB[2] * (1.0 + rng.normal(0.0, 5e-3))
    if extra_quads:
        ks = [0.02, -0.02, 0.015, -0.015]
        res, i, nmid, first_bend = [], 0, 0, True
        while i < len(out):
            e = out[i]
            res.append(e)
            if e.type == el.QUAD and e.name == "QF":
                first_bend = True
            if e.type in (el.BENDMPOLE, el.BEND) and first_bend:
                first_bend = False
                d = out[i + 2]  # bend, its error kick, then the 0.5 m drift
                assert d.type == el.DRIFT and abs(d.len - 0.5) < 1e-12
                fam = nmid % 4
                res += [out[i + 1], el.quad(0.1, ks[fam], f"QM{fam + 1}"), el.drift(0.4, "D3")]
                nmid += 1
                i += 3
                continue
            i += 1
        out = res
    out.append(el.cavity(0.0, 1.5, RF_FREQ, phis, el.DEFAULT_E0, on=rf_on, name="RF"))
    out.append(el.aperture(0.0, (-0.025, 0.025), (-0.012, 0.012), name="AP"))
    return out


def set_rad(elems, flag):
    for e in elems:
        if e.type == el.BENDMPOLE:
            e.rad = flag


def set_cavity(elems, on):
    for e in elems:
        if e.type == el.CAVITY:
            e.on = on


def family(elems, name, etype=None):
    return [i for i, e in enumerate(elems) if e.name == name and (etype is None or e.type == etype)]


def gaussian_beam(n, seed=0, sigmas=BEAM_SIGMAS):
    """n particles, N(0, sigma) per coordinate, Philox-seeded (AoS (n,6))."""
    z = _rng(seed).standard_normal((n, 6))
    return np.ascontiguousarray(z * np.asarray(sigmas)[None, :])


DELTAS = (-3e-2, -1.5e-2, 0.0, 1.5e-2, 3e-2)


def da_grid(nx=101, ny=51, deltas=DELTAS):
    """(x, y) rectangle x in [-20, 20] mm, y in [0, 10] mm at each delta;
    other coordinates 0.  Order: delta, y, x (x fastest)."""
    xs = np.linspace(-0.02, 0.02, nx)
    ys = np.linspace(0.0, 0.01, ny)
    out = np.zeros((len(deltas), ny, nx, 6))
    out[..., 0] = xs[None, None, :]
    out[..., 2] = ys[None, :, None]
    out[..., 5] = np.asarray(deltas)[:, None, None]
    return out.reshape(-1, 6)


def seed_grid_shape(seed):
    """Uneven per-seed grid: N_s ~ U{500..8000} from the seed, n_xy = N_s//5
    points on the DA rectangle as nx x ny with nx ~ 2*ny."""
    ns = int(_rng(10_000 + seed).integers(500, 8001))
    nxy = ns // 5
    ny = max(1, int(round(math.sqrt(nxy / 2.0))))
    nx = max(1, nxy // ny)
    return nx, ny


def batched_workload(nseed=64):
    """Config-2 inputs: per-seed lattices, one AoS beam, offsets[nseed+1]."""
    lats, beams, offs = [], [], [0]
    for s in range(nseed):
        lats.append(da_ring(seed=s, strength_errors=True))
        nx, ny = seed_grid_shape(s)
        g = da_grid(nx, ny)
        beams.append(g)
        offs.append(offs[-1] + len(g))
    return lats, np.ascontiguousarray(np.concatenate(beams)), np.asarray(offs, dtype=np.int64)


def adjoint_knobs(elems):
    """The 8 family knobs of the adjoint workload: QF, QD K; SF, SD
    PolyB[2]; QM1..QM4 K.  Returns [(name, field, index, members)]."""
    out = []
    for name, field, idx in (("QF", el.FIELD_K, 0), ("QD", el.FIELD_K, 0), ("SF", el.FIELD_POLYB, 2),
                             ("SD", el.FIELD_POLYB, 2), ("QM1", el.FIELD_K, 0), ("QM2", el.FIELD_K, 0),
                             ("QM3", el.FIELD_K, 0), ("QM4", el.FIELD_K, 0)):
        t = el.QUAD if field == el.FIELD_K else el.STRMPOLE
        m = family(elems, name, t)
        if m:
            out.append((name, field, idx, m))
    return out


def shift_knob(elems, field, index, members, delta):
    """Copy of elems with the knob moved by delta on every member."""
    out = [e.copy() for e in elems]
    for i in members:
        if field == el.FIELD_K:
            out[i].k += delta
        elif field == el.FIELD_POLYB:
            out[i].polyB[index] += delta
        else:
            out[i].polyA[index] += delta
    return out

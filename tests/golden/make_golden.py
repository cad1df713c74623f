"""Write the golden fixtures of the tracking path from the compiled reference.

Runs in the build container only (it needs oracle/_ref/libaccel_ref.so,
built by `make -C oracle` from the Accel+ sources).  Every fixture is data:
inputs and the reference's outputs for them, written with numpy.savez
(no pickles).  Regenerate with:  python tests/golden/make_golden.py
(or one of the later fixtures alone: python tests/golden/make_golden.py optics2|elements2|transwake|wigtable|naff)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "accel-plus_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyaccp.elements as el  # noqa: E402
import pyaccp.lattices as lat  # noqa: E402
from pyoracle import Reference, beam_stat_ref  # noqa: E402


def probe_particles():
    """21 particles spanning the element maps' branches: three momenta
    (-3%, 0, +3%), both signs of x, y, px; one already lost (x NaN, z = 2);
    two outside the 25 mm x 12 mm aperture."""
    out = []
    i = 0
    for d in (-3e-2, 0.0, 3e-2):
        for x in (-3e-3, 0.0, 2e-3):
            for y in (-1e-3, 1.5e-3):
                out.append([x, 1e-4 * ((i % 3) - 1), y, -5e-5 + 2e-5 * (i % 2), 1e-3 * ((i % 5) - 2), d])
                i += 1
    out.append([np.nan, 1e-4, 1e-3, 0.0, 2.0, 0.0])
    out.append([0.03, 0.0, 1e-3, 0.0, 0.0, 0.0])
    out.append([1e-3, 0.0, -0.02, 0.0, 0.0, 1e-2])
    return np.array(out)


def element_cases():
    """(name, element) pairs, each fixture a single pass of one element."""
    f = lat.RF_FREQ
    c = []
    c.append(("drift", el.drift(1.3)))
    c.append(("quad_f", el.quad(0.3, 1.2)))
    c.append(("quad_d", el.quad(0.3, -1.1)))
    c.append(("quad_k0", el.quad(0.25, 0.0)))
    c.append(("corr_thick", el.corrector(0.2, 1e-4, -2e-4)))
    c.append(("corr_thin", el.corrector(0.0, 1e-4, 1e-4)))
    c.append(("aper_thin", el.aperture(0.0, (-0.025, 0.025), (-0.012, 0.012))))
    c.append(("aper_thick", el.aperture(0.5, (-0.004, 0.0035), (-0.002, 0.002))))
    c.append(("monitor", el.monitor()))
    c.append(("bend_g0", el.bend(1.5, 0.157, 0.05, 0.03, 0.0)))
    c.append(("bend_grad", el.bend(1.2, 0.1, 0.0, 0.0, 0.8)))
    c.append(("bend_defoc", el.bend(1.0, 0.05, 0.02, 0.0, -0.5)))
    c.append(("bmp_dipole", el.bendmpole(1.5, 0.15707963267948966, polyB=[0.0])))
    c.append(("bmp_multi", el.bendmpole(1.5, 0.157, 0.04, 0.02, polyB=[1e-3, 0.3, 5.0, 20.0])))
    c.append(("bmp_dipole_rad", el.bendmpole(1.5, 0.15707963267948966, polyB=[0.0], rad=1)))
    c.append(("bmp_multi_rad", el.bendmpole(1.5, 0.157, 0.04, 0.02, polyB=[1e-3, 0.3, 5.0, 20.0], rad=1)))
    c.append(("bmp_ns4_e2000", el.bendmpole(1.0, 0.12, 0.0, 0.01, polyB=[0.0, 0.2], nslice=4,
                                            energy=2000.0, rad=1)))
    c.append(("smp_sext", el.strmpole(0.2, [0.0, 0.0, 20.0])))
    c.append(("smp_skew", el.strmpole(0.3, [1e-4, 0.5, -8.0, 30.0, 100.0], [2e-4, 0.1, 3.0, -10.0, 50.0])))
    c.append(("smp_thin9", el.strmpole(0.0, [0, 0, 3.0, -40.0, 5e2, 1e4, -3e5, 2e6, 5e7, -1e9],
                                       [0, 0, -1.0, 20.0, 1e3, -2e4, 1e5, 3e6, -2e7, 4e8], max_npole=10)))
    c.append(("smp_ns4", el.strmpole(0.25, [0.0, -0.7, 15.0], nslice=4)))
    c.append(("cav_thin", el.cavity(0.0, 1.5, f, 0.0, 3000.0, on=1)))
    c.append(("cav_thick", el.cavity(0.4, 1.2, 500e6, 0.3, 3000.0, on=1)))
    c.append(("cav_off", el.cavity(0.4, 1.2, 500e6, 0.3, 3000.0, on=0)))
    return c


def corrected_fodo():
    """The FODO ring with its chromaticity corrected to about (2, 1) by two
    correct_chrom steps of the oracle (bit-exact with the reference's)."""
    from pyoracle import Oracle
    O = Oracle()
    arr = el.to_array(lat.fodo_ring())
    sf, sd = lat.family(lat.fodo_ring(), "SF"), lat.family(lat.fodo_ring(), "SD")
    for _ in range(2):
        O.correct(arr, 2, sf, sd, 2.0, 1.0)
    return [arr[i] for i in range(len(arr))]


def optics2(R, out):
    """Twiss functions, second-order chromaticity, 6D closed orbit and
    momentum aperture (aline.cpp:781-889, :1044-1158).  getco4 has no
    fixture: the reference's returns the one-turn image of the off-momentum
    origin, not a closed orbit (its 4x4 block is uninitialised memory)."""
    fodo = lat.fodo_ring()
    idx = sorted(lat.family(fodo, "QF") + lat.family(fodo, "SD"))
    line = R.load(fodo)
    tw0 = line.calctwiss0(0.0)
    tw = line.calctwiss(idx, 1e-4)
    cf = corrected_fodo()
    chrom2 = R.load(cf).calcchrom2()
    b2 = np.array([e.polyB[2] for e in cf if e.type == el.STRMPOLE])
    da = lat.da_ring(0)
    l6 = R.load(da)
    phis = l6.setrfphase()
    co6 = l6.getco6()
    locs = lat.family(da, "QF")[:4]
    lm = R.load(da)
    lm.setrfphase()
    lm.set_rad(1)
    lm.set_cavity(1)
    ma = lm.trackma(0.01, 0.0005, 64, 19, locs)
    np.savez_compressed(os.path.join(out, "optics2.npz"), twiss0=tw0, twiss_idx=np.array(idx, dtype=np.int32),
                        twiss=tw, chrom2_b2=b2, chrom2=chrom2, phis=np.array([phis]), co6=co6,
                        ma_locs=np.array(locs, dtype=np.int32), ma_avg=np.array([ma[0]]), ma_plus=ma[1],
                        ma_minus=ma[2])
    print("optics2:", tw0, chrom2, co6, ma)


def elements2_cases():
    """Matrix, SyncRad (damping only: the excitation draws from rand()),
    CrabCav and RFMult (no noise), ideal wigglers (horizontal, vertical, with
    classical radiation), plus a BendMPole under a lumped-radiation flag (the
    bend then runs without radiation)."""
    rng = np.random.default_rng(7)
    m = np.eye(6) + 0.05 * rng.standard_normal((6, 6))
    bend = el.bendmpole(1.5, 2 * np.pi / 40, polyB=[0, 0.1, 2.0], rad=el.RAD_LUMP, name="BL")
    return [
        ("matrix_thin", el.matrix(0.0, m.ravel())),
        ("matrix_thick", el.matrix(0.5, m.ravel())),
        ("syncrad_damp", el.syncrad(0.0, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0,
                                    rad=el.RAD_LUMP_DAMP)),
        ("syncrad_damp_thick", el.syncrad(0.4, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0,
                                          rad=el.RAD_LUMP_DAMP)),
        ("crab", el.crabcav(0.0, 1.5, 0.5, 1.5e9, 0.3, 3000.0)),
        ("crab_thick", el.crabcav(0.3, 1.5, 0.5, 1.5e9, 0.3, 3000.0)),
        ("bend_lump", bend),
        ("rfmult", el.rfmult(0.0, 5e8, 0.2, [1e-4, 2e-3, 0.5, 3.0], [0.0, 1e-3, 0.2, 0.0])),
        ("rfmult_thick", el.rfmult(0.3, 5e8, 0.2, [1e-4, 2e-3, 0.5], [0.0, 1e-3, 0.2])),
        ("wiggler_h", el.idealwig(2.0, 20, 0.1, 0.5, 0)),
        ("wiggler_v_rad", el.idealwig(2.0, 20, 0.1, 0.5, 1, rad=el.RAD_DAMP)),
        ("wiggler_h_rad2", el.idealwig(1.0, 7, 0.08, 0.3, 0, rad=el.RAD_QE)),
    ]


def elements2(R, out):
    P = probe_particles()
    res = {"particles": P}
    for name, e in elements2_cases():
        res[name] = R.load([e]).elem_pass(0, P)
    np.savez_compressed(os.path.join(out, "elements2.npz"), **res)
    print("elements2:", sorted(res))


def wigtable_particles():
    """The probe particles plus 300 in a cloud over and beyond the tables'
    grids, 20 of them exactly on grid nodes."""
    rng = np.random.default_rng(11)
    c = np.zeros((300, 6))
    c[:, 0] = rng.uniform(-6e-3, 6e-3, 300)
    c[:, 1] = rng.normal(0, 1e-4, 300)
    c[:, 2] = rng.uniform(-3e-3, 3e-3, 300)
    c[:, 3] = rng.normal(0, 5e-5, 300)
    c[:, 4] = rng.normal(0, 3e-3, 300)
    c[:, 5] = rng.normal(0, 1e-2, 300)
    c[:20, 0] = np.linspace(-4e-3, 4e-3, 17)[rng.integers(0, 17, 20)]
    c[:20, 2] = np.linspace(-2e-3, 2e-3, 9)[rng.integers(0, 9, 20)]
    return np.concatenate([probe_particles(), c])


def wigtable_cases():
    """Thin and thick tables on uniform grids, a table on descending grids,
    and the largest table (101 x 21) on non-uniform grids."""
    rng = np.random.default_rng(13)

    def tables(xl, yl, a):
        X, Y = np.meshgrid(xl, yl)
        xk = a * (np.sin(300 * X) * np.cosh(400 * Y) + 0.1 * rng.standard_normal(X.shape))
        yk = a * (np.cos(250 * X) * np.sinh(500 * Y) + 0.1 * rng.standard_normal(X.shape))
        return xk, yk
    xl, yl = np.linspace(-4e-3, 4e-3, 17), np.linspace(-2e-3, 2e-3, 9)
    xk, yk = tables(xl, yl, 2e-5)
    xn = np.sort(np.concatenate([[-5e-3, 5e-3], rng.uniform(-5e-3, 5e-3, 99)]))
    yn = np.sort(np.concatenate([[-2.5e-3, 2.5e-3], rng.uniform(-2.5e-3, 2.5e-3, 19)]))
    xkn, ykn = tables(xn, yn, 1e-5)
    return [
        ("wt_thin", el.wigtable(0.0, 1, xl, yl, xk, yk)),
        ("wt_thick", el.wigtable(1.5, 10, xl, yl, xk, yk)),
        ("wt_desc", el.wigtable(0.0, 1, xl[::-1], yl[::-1], xk[::-1, ::-1], yk[::-1, ::-1])),
        ("wt_max", el.wigtable(0.8, 4, xn, yn, xkn, ykn)),
    ]


def wigtable(R, out):
    P = wigtable_particles()
    res = {"particles": P}
    for name, e in wigtable_cases():
        res[name] = R.load([e]).elem_pass(0, P)
    np.savez_compressed(os.path.join(out, "wigtable.npz"), **res)
    print("wigtable:", sorted(res))


def naff_signals():
    """Signals for naff: (data, xlow, xhi) with lines anywhere in (0, 1/2),
    two lines of similar strength, noise, a constant, a line at the lowest
    bins, a restricted window, and a NaN sample."""
    rng = np.random.default_rng(17)
    out = []
    for n in (16, 64, 256, 1024, 4096):
        for _ in range(4):
            nu = rng.uniform(0.02, 0.48)
            t = np.arange(n)
            x = np.cos(2 * np.pi * nu * t + rng.uniform(0, 6)) + 0.3 * np.sin(2 * np.pi * rng.uniform(0, 0.5) * t)
            out.append((x + 1e-3 * rng.standard_normal(n), 0.0, 0.5))
    t = np.arange(512)
    two = np.cos(2 * np.pi * 0.21 * t) + 0.98 * np.cos(2 * np.pi * 0.3 * t)
    out.append((two, 0.0, 0.5))
    out.append((two, 0.25, 0.35))
    out.append((np.full(128, 0.7), 0.0, 0.5))
    out.append((np.cos(2 * np.pi * 1.3 / 256 * np.arange(256)), 0.0, 0.5))
    out.append((np.cos(2 * np.pi * 0.4999 * np.arange(256)), 0.0, 0.5))
    noisy = rng.standard_normal(256)
    noisy[100] = np.nan
    out.append((noisy, 0.1, 0.4))
    return out


def naff(R, out):
    """naff on synthetic signals; tracktune, trackDiffusion and trackdnudexy
    on the FODO ring and on a DA ring with errors (radiation and cavities
    off, as TRACK_DNUDJ sets them, process.cpp:2619-2621)."""
    sig = naff_signals()
    res = {"naff": np.array([R.naff(d, lo, hi) for d, lo, hi in sig])}
    rings = {"fodo": lat.fodo_ring(), "da": lat.da_ring(2)}
    for name, ring in rings.items():
        line = R.load(ring)
        line.set_rad(0)
        line.set_cavity(0)
        res[f"tune_{name}"] = np.array([line.tracktune(300, 1e-3, 5e-4), line.tracktune(128, 4e-3, 2e-3, 0.2, 0.35,
                                                                                          0.2, 0.3)])
        res[f"dnudexy_{name}"] = line.trackdnudexy(256, 0.004, 0.003)
        x0 = np.array([1e-4, 2e-3, 5e-3, 8e-3, 1.2e-2, 0.05])
        y0 = np.array([1e-4, 1e-3, 3e-3, 2e-3, 4e-3, 0.02])
        tot, nux, nuy, dif = line.trackdiffusion(512, x0, y0, 0.05, 0.45, 0.05, 0.45)
        res[f"diffusion_{name}"] = np.array([nux, nuy, dif])
        res[f"diffusion_total_{name}"] = np.array([tot])
    np.savez_compressed(os.path.join(out, "naff.npz"), **res)
    print("naff:", {k: v.shape for k, v in res.items()})


def transwake_case():
    """A 2 nC bunch of 1000 particles (one lost, a few sharing z) through a
    transverse resonator wake (x: 1e6, 1.5 GHz, Q 5; y: 2e6, 2 GHz, Q 3)."""
    b = lat.gaussian_beam(1000, seed=5)
    b[7, 0] = np.nan
    b[10:14, 4] = b[9, 4]
    return b, el.transwake(0.0, 2.0, 1.0e6, 1.5, 5.0, 2.0e6, 2.0, 3.0, on=1)


def transwake(R, out):
    """ATransWake on one thread (the whole beam is its slice)."""
    b, w = transwake_case()
    line = R.load([w])
    line.set_threads(1)
    np.savez_compressed(os.path.join(out, "transwake.npz"), beam=b, pass1=line.track(b, 1))
    print("transwake written")


def main():
    R = Reference()
    if sys.argv[1:] == ["transwake"]:
        transwake(R, HERE)
        return
    if sys.argv[1:] == ["naff"]:
        naff(R, HERE)
        return
    if sys.argv[1:] == ["wigtable"]:
        wigtable(R, HERE)
        return
    if sys.argv[1:] == ["optics2"]:
        optics2(R, HERE)
        return
    if sys.argv[1:] == ["elements2"]:
        elements2(R, HERE)
        return
    out = os.path.join(HERE)
    P = probe_particles()

    # 1. one pass of each element
    res = {"particles": P}
    for name, e in element_cases():
        line = R.load([e])
        res[name] = line.elem_pass(0, P)
    np.savez_compressed(os.path.join(out, "elements.npz"), **res)

    # 2. config-0 ring, 1000 Gaussian particles, 1 and 10 turns
    ring = lat.fodo_ring()
    beam = lat.gaussian_beam(1000, seed=0)
    line = R.load(ring)
    np.savez_compressed(os.path.join(out, "ring_turns.npz"), beam=beam,
                        turn1=line.track(beam, 1), turn10=line.track(beam, 10))

    # 3. DA lattice (seed-0 errors, cavity on at phis = 0, radiation off):
    #    21 x 11 x 5 grid, 256 turns
    da = lat.da_ring(0)
    grid = lat.da_grid(21, 11)
    line = R.load(da)
    np.savez_compressed(os.path.join(out, "da_grid.npz"), grid=grid, turn256=line.track(grid, 256))

    # 4. optics by tracking and one correction step of each kind
    line = R.load(ring)
    m0 = line.getm66(0.0, 1e-8)
    m1 = line.getm66(1e-3, 1e-8)
    tune = np.array(line.tune())
    chrom = np.array(line.chrom())
    rt = line.resp(1, "Quad", "QF", "QD")
    rc = line.resp(2, "StrMPole", "SF", "SD")
    ct = np.array(line.correct(1, "Quad", "QF", "QD", 0.28, 0.24))
    kqf, kqd = line.getfield(1, "K"), line.getfield(7, "K")
    line2 = R.load(ring)
    cc = np.array(line2.correct(2, "StrMPole", "SF", "SD", 2.0, 1.0))
    b2 = np.array([line2.get_mpole(3, "PolynomB", 2), line2.get_mpole(9, "PolynomB", 2)])
    line3 = R.load(da)
    line3.set_cavity(0)
    tune_da = np.array(line3.tune())
    np.savez_compressed(os.path.join(out, "optics.npz"), m66=m0, m66_dpp=m1, tune=tune, chrom=chrom,
                        resp_tune=rt, resp_chrom=rc, corrected_tune=ct, k_after=np.array([kqf, kqd]),
                        corrected_chrom=cc, b2_after=b2, tune_da=tune_da)

    # 5. radiation: one turn with classical radiation; RF phase of the DA ring
    rad_ring = lat.fodo_ring()
    lat.set_rad(rad_ring, 1)
    line = R.load(rad_ring)
    rad1 = line.track(beam[:200], 1)
    line = R.load(da)
    phis = line.setrfphase()
    np.savez_compressed(os.path.join(out, "radiation.npz"), beam=beam[:200], turn1=rad1,
                        phis=np.array([phis]))

    # 6. trackDA as the DAMA block runs it: phase from setrfphase, then
    #    radiation (damping) and cavities on (process.cpp:2485-2486)
    line = R.load(da)
    line.setrfphase()
    line.set_rad(1)
    line.set_cavity(1)
    area, rx, ry = line.trackda(0.02, 0.001, 0.01, 64, 11, 10)
    line = R.load(da)
    line.setrfphase()
    line.set_rad(1)
    line.set_cavity(1)
    area_w, rxw, ryw = line.trackda(0.02, 0.001, 0.01, 64, 11, 10, negweight=True)
    np.savez_compressed(os.path.join(out, "trackda.npz"), area=np.array([area, area_w]), rx=rx, ry=ry)

    # 7. beam statistics with lost particles
    b = beam[:50].copy()
    b[3, 0] = np.nan
    b[17, 0] = np.nan
    nl, ar, sg, sr = beam_stat_ref(R, b)
    np.savez_compressed(os.path.join(out, "beamstat.npz"), beam=b, nlive=np.array([nl]), mean=ar,
                        sigma=sg, cross=sr)
    optics2(R, out)
    elements2(R, out)
    transwake(R, out)
    wigtable(R, out)
    naff(R, out)
    print("fixtures written to", out)


if __name__ == "__main__":
    main()

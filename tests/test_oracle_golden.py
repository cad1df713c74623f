"""The oracle against the reference: bit-exact on every committed fixture
(written by tests/golden/make_golden.py from the compiled reference), and
against the live reference library where it is built."""
import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from conftest import golden
from helpers import same_bits

make_golden = pytest.importorskip("golden.make_golden")


def test_element_fixtures_bit_exact(oracle):
    g = golden("elements.npz")
    P = g["particles"]
    for name, e in make_golden.element_cases():
        out = oracle.track([e], P, 1)
        assert same_bits(out, g[name]), name


def test_ring_turns_bit_exact(oracle):
    g = golden("ring_turns.npz")
    ring = lat.fodo_ring()
    assert same_bits(g["beam"], lat.gaussian_beam(1000, seed=0))
    assert same_bits(oracle.track(ring, g["beam"], 1), g["turn1"])
    assert same_bits(oracle.track(ring, g["beam"], 10), g["turn10"])


def test_da_grid_bit_exact(oracle):
    g = golden("da_grid.npz")
    out = oracle.track(lat.da_ring(0), g["grid"], 256)
    assert same_bits(out, g["turn256"])
    lost = np.isnan(out[:, 0])
    codes = out[lost, 4]
    # aperture codes (accel.cpp:352), or NaN for orbits that overflowed
    # before reaching the aperture
    finite = codes[~np.isnan(codes)]
    assert set(np.unique(finite)).issubset({1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0})


def test_optics_bit_exact(oracle):
    g = golden("optics.npz")
    ring = lat.fodo_ring()
    assert same_bits(oracle.getm66(ring, 0.0, 1e-8), g["m66"])
    assert same_bits(oracle.getm66(ring, 1e-3, 1e-8), g["m66_dpp"])
    assert same_bits(oracle.tune(ring), g["tune"])
    assert same_bits(oracle.chrom(ring), g["chrom"])
    arr = el.to_array(ring)
    qf, qd = lat.family(ring, "QF"), lat.family(ring, "QD")
    sf, sd = lat.family(ring, "SF"), lat.family(ring, "SD")
    assert same_bits(oracle.resp(arr, 1, qf, qd), g["resp_tune"])
    assert same_bits(oracle.resp(arr, 2, sf, sd), g["resp_chrom"])
    ct = oracle.correct(arr, 1, qf, qd, 0.28, 0.24)
    assert same_bits(ct, g["corrected_tune"])
    assert same_bits([arr[qf[0]].k, arr[qd[0]].k], g["k_after"])
    arr2 = el.to_array(lat.fodo_ring())
    cc = oracle.correct(arr2, 2, sf, sd, 2.0, 1.0)
    assert same_bits(cc, g["corrected_chrom"])
    assert same_bits([arr2[sf[0]].polyB[2], arr2[sd[0]].polyB[2]], g["b2_after"])
    da = lat.da_ring(0)
    lat.set_cavity(da, 0)
    assert same_bits(oracle.tune(da), g["tune_da"])


def test_radiation_and_rf_phase_bit_exact(oracle):
    g = golden("radiation.npz")
    ring = lat.fodo_ring()
    lat.set_rad(ring, 1)
    assert same_bits(oracle.track(ring, g["beam"], 1), g["turn1"])
    arr = el.to_array(lat.da_ring(0))
    assert same_bits([oracle.setrfphase(arr)], g["phis"])
    # the reference leaves radiation and cavities as setrfphase sets them
    assert all(arr[i].rad == 0 for i in range(len(arr)) if arr[i].type == el.BENDMPOLE)
    # a LongWake sits out the probe turn and keeps its flag
    wk = el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.15, 0.15, energy=3000.0, on=1)
    arrw = el.to_array(lat.da_ring(0) + [wk])
    assert same_bits([oracle.setrfphase(arrw)], g["phis"]) and arrw[len(arrw) - 1].on == 1


def test_trackda_bit_exact(oracle):
    g = golden("trackda.npz")
    for k, neg in enumerate((False, True)):
        arr = el.to_array(lat.da_ring(0))
        oracle.setrfphase(arr)
        for i in range(len(arr)):
            if arr[i].type == el.BENDMPOLE:
                arr[i].rad = 1
            if arr[i].type == el.CAVITY:
                arr[i].on = 1
        area, rx, ry = oracle.trackda(arr, 0.02, 0.001, 0.01, 64, 11, 10, negweight=neg)
        assert same_bits([area], g["area"][k:k + 1])
        if not neg:
            assert same_bits(rx, g["rx"]) and same_bits(ry, g["ry"])


def test_beam_stat_bit_exact(oracle):
    g = golden("beamstat.npz")
    s = oracle.beam_stat(g["beam"])
    assert s[1] == g["nlive"][0]
    assert same_bits(s[2:8], g["mean"])
    assert same_bits(s[8:14], g["sigma"])
    assert same_bits(s[14:20], g["cross"])


def test_oracle_matches_live_reference(oracle, reference):
    """Fresh inputs the fixtures do not hold: the DA lattice of seed 3 with
    strength errors, radiation on, 5 turns."""
    da = lat.da_ring(3, strength_errors=True, extra_quads=True)
    lat.set_rad(da, 1)
    beam = lat.gaussian_beam(300, seed=7)
    line = reference.load(da)
    assert same_bits(oracle.track(da, beam, 5), line.track(beam, 5))
    ring = lat.fodo_ring()
    line = reference.load(ring)
    # ALine::pass(r, N, st, ed): one pass over an element range
    assert same_bits(oracle.track(ring, beam, 1, 17, 100), line.track_range(beam, 17, 100))


def test_optics_services_oracle(oracle):
    """Twiss functions, second-order chromaticity, 6D closed orbit and
    momentum aperture against the reference (optics2.npz).  Twiss and the
    chromaticity fit are bit-exact; the closed orbit solves (I - M) x = r1
    by elimination where the reference inverts by SVD, so it agrees to
    round-off; the momentum aperture follows bit-exact tracking."""
    from tests.golden.make_golden import corrected_fodo
    g = golden("optics2.npz")
    fodo = lat.fodo_ring()
    assert same_bits(oracle.calctwiss0(fodo, 0.0), g["twiss0"])
    assert same_bits(oracle.calctwiss(fodo, g["twiss_idx"], 1e-4), g["twiss"])
    cf = corrected_fodo()
    assert same_bits([e.polyB[2] for e in cf if e.type == el.STRMPOLE], g["chrom2_b2"])
    assert same_bits(oracle.calcchrom2(cf), g["chrom2"])
    arr = el.to_array(lat.da_ring(0))
    assert same_bits([oracle.setrfphase(arr)], g["phis"])
    co6 = oracle.getco6(arr)
    assert np.all(np.abs(co6 - g["co6"]) <= 1e-12 * np.abs(g["co6"]).max())
    arr = el.to_array(lat.da_ring(0))
    oracle.setrfphase(arr)
    for i in range(len(arr)):
        if arr[i].type == el.BENDMPOLE:
            arr[i].rad = 1
        if arr[i].type == el.CAVITY:
            arr[i].on = 1
    avg, p, m = oracle.trackma(arr, 0.01, 0.0005, 64, 19, g["ma_locs"])
    assert avg == g["ma_avg"][0] and same_bits(p, g["ma_plus"]) and same_bits(m, g["ma_minus"])


def test_element_fixtures2_bit_exact(oracle):
    """Matrix, SyncRad (damping), CrabCav and a bend under a lumped-radiation
    flag: one pass, bit for bit with the reference (elements2.npz)."""
    from tests.golden.make_golden import elements2_cases
    g = golden("elements2.npz")
    for name, e in elements2_cases():
        assert same_bits(oracle.track([e], g["particles"], 1), g[name]), name


def test_radiation_equilibrium_oracle_vs_reference():
    """Quantum excitation (flag 2) on the FODO ring at chromaticity +1 with
    the cavity at the synchronous phase: 8192 particles, 10000 turns to
    equilibrium, then 20 records 100 turns apart
    (tests/golden/make_equilibrium_ref.py); energy spread, horizontal size
    and projected emittance averaged over the records.
    * reference_crn: the compiled reference fed the oracle's normals through
      rand() (common random numbers).  The tracks stay together, so the
      equilibria agree to round-off growth: within 0.5% (0.06% measured), and
      losses within 0.25% of the bunch.
    * reference: the reference with its own rand() noise, an independent
      sample of the same equilibrium.  With about 8150 particles a width
      carries 0.8% sampling error per run (1/sqrt(2N)) and the emittance
      ~1.1%, and the 2000-turn window is shorter than the damping times, so
      the records are one sample: two runs differ by 1.1-1.6% at one standard
      error.  Bound: 5% (three standard errors), losses within 1% of the
      bunch."""
    g = golden("equilibrium_ref.npz")
    o = g["oracle"]
    n = int(g["n"][0])
    for run, tol, loss_tol in (("reference_crn", 0.005, 0.0025), ("reference", 0.05, 0.01)):
        r = g[run]
        for k, name in enumerate(("sigma_delta", "sigma_x", "emittance_x")):
            mo, mr = o[:, k].mean(), r[:, k].mean()
            print(f"{run} {name}: oracle {mo:.5e} reference {mr:.5e} ({mo / mr - 1:+.2%})")
            assert abs(mo / mr - 1) < tol, (run, name, mo, mr)
        print(f"{run} live: oracle {o[-1, 3]:.0f} reference {r[-1, 3]:.0f}")
        assert abs(o[-1, 3] - r[-1, 3]) <= loss_tol * n, (run, o[-1, 3], r[-1, 3])


def noise_bound(r, quiet):
    """Per coordinate: 1e-6 of the largest noise displacement |r - quiet|
    over the finite particles.  The harness rounds each Philox uniform u to
    rand()'s grid (relative 5e-10), and Box-Muller's radius sqrt(-2 ln u)
    magnifies that by 1/(2 u ln(1/u)); over the few thousand draws of a case
    the smallest u is ~1e-4, so single normals differ by up to ~1e-7 of
    their size (5e-8 of the noise measured)."""
    fin = np.isfinite(r[:, 0])
    return 1e-6 * np.abs(r - quiet)[fin].max(axis=0), fin


def test_noise_oracle_vs_reference_crn(oracle):
    """CrabCav phase and amplitude jitter (both, each alone), RFMult phase
    jitter, SyncRad excitation, BendMPole quantum excitation and a line of
    three noisy elements: 3 turns from turn 7 with particle ids from 1000
    against the compiled reference fed the same normals in its own draw
    order (noise.npz, tests/golden/make_noise.py).  A normal taken in
    another order or from another (particle, turn, element) moves a
    coordinate by the whole noise, 1e6 times the bound."""
    from tests.golden.make_noise import noise_cases
    g = golden("noise.npz")
    P = g["particles"]
    kw = dict(seed=int(g["seed"][0]), turn0=int(g["turn0"][0]), pid0=int(g["pid0"][0]))
    nt = int(g["turns"][0])
    for name, noisy, quiet in noise_cases():
        r = g[name]
        q = oracle.track(quiet, P, nt, **kw)
        tol, fin = noise_bound(r, q)
        assert tol.max() > 1e-12, name  # the noise is there
        o = oracle.track(noisy, P, nt, **kw)
        assert np.array_equal(np.isfinite(o[:, 0]), fin), name
        err = np.abs(o - r)[fin].max(axis=0)
        print(f"{name}: noise up to {tol.max() * 1e6:.2e}, oracle - reference {err.max():.2e}")
        assert (err <= tol).all(), (name, err, tol)


def test_transwake_bit_exact(oracle):
    """The transverse wake over the whole beam (one reference thread),
    pairwise as the reference sums it: bit for bit."""
    from tests.golden.make_golden import transwake_case
    g = golden("transwake.npz")
    b, w = transwake_case()
    assert same_bits(b, g["beam"])
    out = oracle.track([w], b, 1)
    assert same_bits(out, g["pass1"])
    assert np.nanmax(np.abs(out[:, 1] - b[:, 1])) > 1e-8  # the kick is there


def test_wigtable_fixtures_bit_exact(oracle):
    """WigTable, thin and sliced, ascending, descending and non-uniform
    grids up to the largest table (101 x 21), particles inside, on the nodes
    of and beyond the grid: one pass, bit for bit with the reference
    (wigtable.npz)."""
    from tests.golden.make_golden import wigtable_cases
    g = golden("wigtable.npz")
    for name, e in wigtable_cases():
        assert same_bits(oracle.track([e], g["particles"], 1), g[name]), name


def _quiet(ring):
    """The ring with radiation and cavities off (TRACK_DNUDJ, process.cpp:2619-2621)."""
    arr = el.to_array(ring)
    for i in range(len(arr)):
        if arr[i].type == el.CAVITY:
            arr[i].on = 0
        if arr[i].type in (el.BENDMPOLE, el.IDEALWIG):
            arr[i].rad = 0
    return arr


def test_naff_and_tune_diagnostics_bit_exact(oracle):
    """naff (its own transform, the reference's golden-section steps) on
    synthetic signals, and tracktune, trackdnudexy, trackDiffusion on two
    rings: bit for bit with the reference (naff.npz)."""
    from tests.golden.make_golden import naff_signals
    g = golden("naff.npz")
    got = np.array([oracle.naff(d, lo, hi) for d, lo, hi in naff_signals()])
    assert same_bits(got, g["naff"])
    for name, ring in (("fodo", lat.fodo_ring()), ("da", lat.da_ring(2))):
        arr = _quiet(ring)
        t = np.array([oracle.tracktune(arr, 300, 1e-3, 5e-4), oracle.tracktune(arr, 128, 4e-3, 2e-3, 0.2, 0.35,
                                                                                0.2, 0.3)])
        assert same_bits(t, g[f"tune_{name}"]), name
        assert same_bits(oracle.trackdnudexy(arr, 256, 0.004, 0.003), g[f"dnudexy_{name}"]), name
        x0 = np.array([1e-4, 2e-3, 5e-3, 8e-3, 1.2e-2, 0.05])
        y0 = np.array([1e-4, 1e-3, 3e-3, 2e-3, 4e-3, 0.02])
        tot, nux, nuy, dif = oracle.trackdiffusion(arr, 512, x0, y0, 0.05, 0.45, 0.05, 0.45)
        assert same_bits(np.array([nux, nuy, dif]), g[f"diffusion_{name}"]), name
        assert same_bits(np.array([tot]), g[f"diffusion_total_{name}"]), name
        assert np.isnan(dif[-1]) and not np.isnan(dif[:2]).any()  # small amplitudes survive, 5 cm is lost

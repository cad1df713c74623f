"""Tune diagnostics on the GPU: naff, tracktune, trackdnudexy and
trackDiffusion (utility.cpp:21-96; aline.cpp:716-779, :1159-1186,
:1421-1484) against the compiled reference's outputs (naff.npz).

The refined frequency is the minimum of a golden-section search that stops
when its bracket is below 3e-8 (|x1| + |x2|); the GPU evaluates the same
projections with other round-off (its sincos, a parallel sum), so late
comparisons of two nearly equal projections can go the other way.  Tunes
therefore agree to the search tolerance, 6e-8 of the tune, not bit for bit.
"""
import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from conftest import golden

pytestmark = pytest.mark.gpu

TUNE_TOL = 6e-8


def _quiet(ring):
    out = []
    for e in ring:
        e = e.copy()
        if e.type == el.CAVITY:
            e.on = 0
        if e.type in (el.BENDMPOLE, el.IDEALWIG):
            e.rad = 0
        out.append(e)
    return out


def _close_tunes(got, ref, what):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    assert np.array_equal(np.isnan(got), np.isnan(ref)), what
    ok = ~np.isnan(ref)
    err = np.abs(got[ok] - ref[ok])
    print(what, "max tune difference", err.max() if err.size else 0.0)
    assert np.all(err <= TUNE_TOL * np.abs(ref[ok]) + 1e-12), (what, got, ref)


def test_naff_signals_match_reference(gpu_lib):
    """Lines anywhere in (0, 1/2), two close lines, a restricted window and a
    NaN sample (the reference keeps its first bin and shrinks the bracket
    from above; the GPU takes the same steps) to the search tolerance.  The
    three degenerate signals (a constant; a line 1.3 bins from zero, which
    its own image pulls; a line at the Nyquist frequency, which the packed
    spectrum puts in bin 0) leave a projection that is flat to round-off over
    the bracket, so there both codes pick the same bin and end anywhere in
    it: within one bin."""
    from tests.golden.make_golden import naff_signals
    g = golden("naff.npz")
    sig = naff_signals()
    got = np.array([gpu_lib.naff(d, lo, hi)[0] for d, lo, hi in sig])
    ref = g["naff"]
    flat = [22, 23, 24]
    sharp = [i for i in range(len(sig)) if i not in flat]
    _close_tunes(got[sharp], ref[sharp], "naff")
    for i in flat:
        assert abs(got[i] - ref[i]) <= 1.0 / len(sig[i][0]), (i, got[i], ref[i])


def test_naff_recovers_pure_lines(gpu_lib):
    """A closed form: one line of frequency nu (plus a weaker one far away)
    is found to 1e-6 of nu at 1024 samples, in one batch of signals."""
    rng = np.random.default_rng(5)
    nus = rng.uniform(0.05, 0.45, 64)
    t = np.arange(1024)
    sig = np.array([np.cos(2 * np.pi * nu * t + 1.0) + 0.05 * np.cos(2 * np.pi * 0.02 * t) for nu in nus])
    got = gpu_lib.naff(sig, 0.04, 0.5)
    assert np.all(np.abs(got - nus) < 1e-6 * nus), np.abs(got - nus).max()


def test_tracktune_dnudexy_diffusion_match_reference(gpu_lib):
    g = golden("naff.npz")
    for name, ring in (("fodo", lat.fodo_ring()), ("da", lat.da_ring(2))):
        L = gpu_lib.lattice(_quiet(ring))
        t = np.array([L.tracktune(300, 1e-3, 5e-4), L.tracktune(128, 4e-3, 2e-3, 0.2, 0.35, 0.2, 0.3)])
        _close_tunes(t, g[f"tune_{name}"], f"tracktune {name}")
        d = L.trackdnudexy(256, 0.004, 0.003)
        ref = g[f"dnudexy_{name}"]
        print(name, "dnudexy", d, ref, np.abs(d - ref) / np.abs(ref))
        # the twelve tunes agree to round-off here (the searches take the same
        # branches); a flipped late comparison would move a tune by the 6e-8
        # tolerance, 1e-4 of these slopes over amplitudes x^2/betx ~ 1e-6 m
        assert np.allclose(d, ref, rtol=1e-4, atol=0), (d, ref)
        x0 = np.array([1e-4, 2e-3, 5e-3, 8e-3, 1.2e-2, 0.05])
        y0 = np.array([1e-4, 1e-3, 3e-3, 2e-3, 4e-3, 0.02])
        tot, nux, nuy, dif = L.trackdiffusion(512, x0, y0, 0.05, 0.45, 0.05, 0.45)
        rd = g[f"diffusion_{name}"]
        _close_tunes(nux, rd[0], f"diffusion nux {name}")
        _close_tunes(nuy, rd[1], f"diffusion nuy {name}")
        print(name, "dif", dif, rd[2])
        assert np.array_equal(np.isnan(dif), np.isnan(rd[2]))
        live = ~np.isnan(dif)
        # dif = ln(2 |dnu| / nt), nt = 512.  Where the reference's tune change
        # |dnu| is well above the search tolerance (1e-6 against ~2e-8) the
        # two agree to 1e-2; below it both measure the search's own noise,
        # and the GPU's |dnu| must be below 2e-6 as well
        big = live & (np.exp(rd[2]) * 256 > 1e-6)
        assert np.allclose(dif[big], rd[2][big], rtol=0, atol=1e-2), (dif, rd[2])
        assert np.all(np.exp(dif[live & ~big]) * 256 < 2e-6), (dif, rd[2])


def test_frequency_map_in_groups(gpu_lib):
    """A frequency map whose turn-by-turn record exceeds the 8 GB a single
    tracking launch records (30000 points x 8192 turns x 48 B = 11.8 GB) is
    tracked in groups: the result equals two calls on the halves, bit for
    bit (no randomness in this lattice, and each group's particles keep
    their place)."""
    L = gpu_lib.lattice(_quiet(lat.da_ring(0)))
    rng = np.random.default_rng(2)
    x0, y0 = rng.uniform(1e-4, 6e-3, 30000), rng.uniform(1e-4, 3e-3, 30000)
    tot, nux, nuy, dif = L.trackdiffusion(8192, x0, y0, 0.05, 0.45, 0.05, 0.45)
    parts = [L.trackdiffusion(8192, x0[a:b], y0[a:b], 0.05, 0.45, 0.05, 0.45) for a, b in ((0, 15000),
                                                                                      (15000, 30000))]
    for k, arr in ((1, nux), (2, nuy), (3, dif)):
        both = np.concatenate([parts[0][k], parts[1][k]])
        assert np.array_equal(np.isnan(arr), np.isnan(both)) and np.array_equal(np.nan_to_num(arr),
                                                                                np.nan_to_num(both))
    assert abs(tot - (parts[0][0] + parts[1][0])) <= 1e-9 * abs(tot)
    assert (~np.isnan(dif)).sum() > 1000

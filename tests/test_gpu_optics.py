"""ALine's lattice services on the GPU tracker against the reference.

getm66, calctune, calcchrom, the tune and chromaticity responses and one
correction step of correct_tune / correct_chrom (aline.cpp:891-1418), the RF
phase (aline.cpp:465-501) and trackDA (aline.cpp:607-715), each against the
compiled reference's outputs in tests/golden (optics.npz, radiation.npz,
trackda.npz).  The GPU tracks the rays to 1e-12 of the reference, so the
matrix and tunes agree to round-off; derived finite differences lose what
their step cancels, and each tolerance below says how much.  The adjoint
responses differentiate the same getm66-based tune exactly and are checked
against central differences of the GPU's own tunes, and against the
reference's +1e-5 forward difference within its truncation error."""
import numpy as np
import pytest

import pyaccp.abi as abi
import pyaccp.elements as el
import pyaccp.lattices as lat
from conftest import golden

pytestmark = pytest.mark.gpu

TUNE, CHROM, FD, ADJ = abi.Lattice.TUNE, abi.Lattice.CHROM, abi.Lattice.RESP_FD, abi.Lattice.RESP_ADJOINT


def fams(ring, *names):
    return [lat.family(ring, n) for n in names]


def test_matrix_tune_chromaticity(gpu_lib):
    g = golden("optics.npz")
    L = gpu_lib.lattice(lat.fodo_ring())
    m = L.getm66()
    scale = np.abs(g["m66"]).max()
    print("m66 max abs diff / max", np.abs(m - g["m66"]).max() / scale)
    assert np.abs(m - g["m66"]).max() <= 1e-12 * scale
    # dp = 1e-3: rays of 1e-8 on a 1e-3 orbit keep 1e-16 * 1e5 relative
    m = L.getm66(dp=1e-3)
    assert np.abs(m - g["m66_dpp"]).max() <= 1e-9 * np.abs(g["m66_dpp"]).max()
    assert np.allclose(L.calctune(), g["tune"], rtol=0, atol=1e-13)
    # a tune difference over 2e-5: 1e-13 / 2e-5 absolute
    assert np.allclose(L.calcchrom(), g["chrom"], rtol=0, atol=1e-7)
    da = lat.da_ring(0)
    lat.set_cavity(da, 0)
    assert np.allclose(gpu_lib.lattice(da).calctune(), g["tune_da"], rtol=0, atol=1e-13)


def test_responses_fd_match_reference(gpu_lib):
    g = golden("optics.npz")
    ring = lat.fodo_ring()
    L = gpu_lib.lattice(ring)
    qf, qd, sf, sd = fams(ring, "QF", "QD", "SF", "SD")
    rt = L.calcresp(TUNE, qf, qd, FD)
    rc = L.calcresp(CHROM, sf, sd, FD)
    print("tune resp rel", np.abs(rt - g["resp_tune"]).max() / np.abs(g["resp_tune"]).max(),
          "chrom resp rel", np.abs(rc - g["resp_chrom"]).max() / np.abs(g["resp_chrom"]).max())
    # tune difference over 1e-5: 1e-13 / 1e-5 = 1e-8 absolute
    assert np.abs(rt - g["resp_tune"]).max() <= 1e-7
    # chromaticity (itself a 2e-5 difference) difference over 1e-5
    assert np.abs(rc - g["resp_chrom"]).max() <= 1e-2 * np.abs(g["resp_chrom"]).max()
    # the responses leave the lattice as they found it
    assert bytes(L.get()) == bytes(el.to_array(ring))


def test_correction_steps_match_reference(gpu_lib):
    g = golden("optics.npz")
    ring = lat.fodo_ring()
    qf, qd, sf, sd = fams(ring, "QF", "QD", "SF", "SD")
    L = gpu_lib.lattice(ring)
    reached = L.correct(TUNE, qf, qd, [0.28, 0.24], FD)
    arr = L.get()
    print("tunes", reached - g["corrected_tune"], "K", [arr[qf[0]].k, arr[qd[0]].k] - g["k_after"])
    assert np.allclose(reached, g["corrected_tune"], rtol=0, atol=1e-9)
    assert np.allclose([arr[qf[0]].k, arr[qd[0]].k], g["k_after"], rtol=1e-8, atol=0)
    assert all(arr[i].k == arr[qf[0]].k for i in qf)
    L = gpu_lib.lattice(lat.fodo_ring())
    start = L.calcchrom()
    reached = L.correct(CHROM, sf, sd, [2.0, 1.0], FD)
    arr = L.get()
    b2 = np.array([arr[sf[0]].polyB[2], arr[sd[0]].polyB[2]])
    print("chrom", start, reached - g["corrected_chrom"], "B2", b2 - g["b2_after"])
    # The reference's chromaticity response is a 1e-5 difference of
    # chromaticities that are themselves 2e-5 differences of tunes from
    # 1e-8 rays: 1e-13 of tune round-off is 1e-3 of the response, and the
    # bends and sextupoles sum their terms in another order than the
    # reference's, so the two codes' responses differ by up to about 1%; the
    # one step from +100 to 2 carries that over |dxi|.  Measured on the
    # oracle (the reference's order, which reaches the fixture's 2.178
    # exactly): one ulp more of SF's PolyB[2] moves the reached xi_x by
    # 0.008, and the sextupole drifts between slices taken as one drift of
    # twice the length, as the kernels take them (the same map in exact
    # arithmetic), by 0.17 = 1.7e-3 of the span; the bound is 4e-3 of it.
    span = np.abs(np.array([2.0, 1.0]) - start).max()
    assert np.allclose(reached, g["corrected_chrom"], rtol=0, atol=4e-3 * span)
    assert np.allclose(b2, g["b2_after"], rtol=3e-2, atol=0)


def test_adjoint_responses(gpu_lib):
    """Adjoint responses against fourth-order central differences of the
    GPU's own tunes and chromaticities (1e-6) and the reference's forward
    differences (their truncation and round-off).
    The knob steps balance truncation against round-off: tunes carry 1e-13
    and a 1e-4 relative K step keeps that below 1e-8; chromaticities carry
    5e-9 and are nearly linear in PolyB[2], so a 1e-2 relative step."""
    g = golden("optics.npz")
    ring = lat.fodo_ring()
    qf, qd, sf, sd = fams(ring, "QF", "QD", "SF", "SD")
    L = gpu_lib.lattice(ring)
    for quantity, f1, f2, ref, field, idx in ((TUNE, qf, qd, g["resp_tune"], el.FIELD_K, 0),
                                              (CHROM, sf, sd, g["resp_chrom"], el.FIELD_POLYB, 2)):
        adj = L.calcresp(quantity, f1, f2, ADJ)
        measure = L.calctune if quantity == TUNE else L.calcchrom
        for k, fam in enumerate((f1, f2)):
            base = ring[fam[0]].k if field == el.FIELD_K else ring[fam[0]].polyB[idx]
            h = (1e-4 if quantity == TUNE else 1e-2) * abs(base)
            v = {}
            for s in (-2, -1, 1, 2):
                L.update(lat.shift_knob(ring, field, idx, fam, s * h))
                v[s] = measure()
            L.update(ring)
            fd = (-v[2] + 8 * v[1] - 8 * v[-1] + v[-2]) / (12 * h)
            print(quantity, k, "adjoint", adj[k], "fd", fd, "reference fd", ref[k])
            assert np.all(np.abs(adj[k] - fd) <= 1e-6 * np.abs(fd).max())
            # the reference's forward differences: truncation 1e-5 for tunes;
            # for chromaticities its ray round-off over the two small steps
            # reaches 6e-3 of the smaller entries
            rtol = 1e-4 if quantity == TUNE else 1e-2
            assert np.all(np.abs(adj[k] - ref[k]) <= rtol * np.abs(ref[k]).max())


def test_adjoint_corrections_reach_targets(gpu_lib):
    ring = lat.fodo_ring()
    qf, qd, sf, sd = fams(ring, "QF", "QD", "SF", "SD")
    L = gpu_lib.lattice(ring)
    for _ in range(3):
        reached = L.correct(TUNE, qf, qd, [0.28, 0.24], ADJ)
    assert np.allclose(reached, [0.28, 0.24], rtol=0, atol=1e-10)
    for _ in range(3):
        reached = L.correct(CHROM, sf, sd, [2.0, 1.0], ADJ)
    assert np.allclose(reached, [2.0, 1.0], rtol=0, atol=1e-6)
    assert np.allclose(L.calctune(), [0.28, 0.24], rtol=0, atol=1e-10)  # the sextupoles keep the tunes


def test_responses_reject_bad_families(gpu_lib):
    ring = lat.fodo_ring()
    L = gpu_lib.lattice(ring)
    b = [i for i, e in enumerate(ring) if e.type == el.BENDMPOLE][:1]
    with pytest.raises(abi.AccpError, match="neither a Quad nor a StrMPole"):
        L.calcresp(TUNE, b, lat.family(ring, "QD"))
    with pytest.raises(abi.AccpError, match="not a StrMPole"):
        L.calcresp(CHROM, lat.family(ring, "QF"), lat.family(ring, "SD"))
    with pytest.raises(abi.AccpError):
        L.calcresp(TUNE, [], lat.family(ring, "QD"))


def test_setrfphase_matches_reference(gpu_lib):
    g = golden("radiation.npz")
    L = gpu_lib.lattice(lat.da_ring(0))
    phis = L.setrfphase()
    assert abs(phis - g["phis"][0]) <= 1e-12 * abs(g["phis"][0])
    arr = L.get()
    # like the reference: radiation off and the cavities left off
    assert all(e.rad == 0 for e in arr if e.type == el.BENDMPOLE)
    assert all(e.on == 0 and e.phis == phis for e in arr if e.type == el.CAVITY)
    assert gpu_lib.lattice(lat.fodo_ring()).setrfphase() == 0.0  # no cavity: nothing to set
    # a LongWake sits out the probe turn (a lone particle would see the whole
    # bunch's own field) and keeps its flag
    wk = el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.15, 0.15, energy=3000.0, on=1)
    Lw = gpu_lib.lattice(lat.da_ring(0) + [wk])
    assert Lw.setrfphase() == phis
    assert [e.on for e in Lw.get() if e.type == el.LONGWAKE] == [1]


def da_lattice(oracle):
    arr = el.to_array(lat.da_ring(0))
    oracle.setrfphase(arr)
    for i in range(len(arr)):
        if arr[i].type == el.BENDMPOLE:
            arr[i].rad = 1
        if arr[i].type == el.CAVITY:
            arr[i].on = 1
    return arr


def test_trackda_matches_reference(gpu_lib, oracle):
    """DA rays: the boundary may move by one step where a point sits on the
    edge of survival (the maps agree only to 1e-12 per turn); with no such
    point the profile and area agree to round-off."""
    g = golden("trackda.npz")
    xmax, xmin, ymax, nstep = 0.02, 0.001, 0.01, 10
    step = (xmax - xmin) / (nstep - 1)
    for k, neg in enumerate((False, True)):
        L = gpu_lib.lattice(da_lattice(oracle))
        area, rx, ry, lt = L.trackda(xmax, xmin, ymax, 64, 11, nstep, negweight=neg, loss_turn=True)
        assert len(rx) == 11 and lt.shape == (110,)
        assert np.all((lt >= 0) & (lt <= 64))
        if not neg:
            rad = np.hypot(rx, ry * xmax / ymax)
            rad_ref = np.hypot(g["rx"], g["ry"] * xmax / ymax)
            moved = np.abs(rad - rad_ref)
            print("rays moved", moved / step)
            assert np.all(moved <= step * (1 + 1e-9))
            if not moved.any():
                assert np.allclose(rx, g["rx"], rtol=1e-12, atol=1e-15)
        if np.array_equal(rx, g["rx"]) and np.array_equal(ry, g["ry"]):
            assert abs(area - g["area"][k]) <= 1e-12 * abs(g["area"][k])
        # loss turns agree with the survival map: the first lost point of a ray
        # is lost before the end, every point before it survives
        lt = lt.reshape(11, nstep)
        for i in range(11):
            lost = np.nonzero(lt[i] < 64)[0]
            if len(lost):
                assert np.all(lt[i, :lost[0]] == 64)


def test_trackda_argument_checks(gpu_lib, oracle):
    L = gpu_lib.lattice(da_lattice(oracle))
    with pytest.raises(abi.AccpError, match="positive"):
        L.trackda(0.02, 0.0, 0.01, 8, 11, 10)
    with pytest.raises(abi.AccpError, match="greater"):
        L.trackda(0.001, 0.02, 0.01, 8, 11, 10)
    area, rx, ry = L.trackda(0.02, 0.001, 0.01, 4, 99, 1)  # clamped to 51 rays, 2 steps
    assert len(rx) == 51


def test_twiss_and_second_order_chromaticity(gpu_lib):
    """Against the reference (optics2.npz): Twiss functions from the
    one-turn matrix to 1e-11 on momentum.  At dp = 1e-4 the 1e-8 rays of
    getm66 ride on a 1.5e-4 m dispersion orbit, so the round-off of some 3000
    sub-steps per turn (summed in another order than the reference's in the
    bends) enters the difference quotient amplified 1e4-fold: 1e-9 there.
    The second-order chromaticity is the slope of chromaticities taken at
    dpp up to 6e-3, where the 1e-8 rays ride on orbits of 6e-3 m: each
    chromaticity keeps eps * 6e-3 / 1e-8 / 2e-5 ~ 3e-6 of round-off, and the
    slope over a 1e-2 span carries ~1e-4 of it."""
    from tests.golden.make_golden import corrected_fodo
    g = golden("optics2.npz")
    L = gpu_lib.lattice(lat.fodo_ring())
    assert np.allclose(L.calctwiss(), g["twiss0"], rtol=1e-11, atol=1e-13)
    tw = L.calctwiss(g["twiss_idx"], dp=1e-4)
    print("twiss at dp=1e-4: max relative difference", float(np.max(np.abs(tw - g["twiss"]) /
                                                                   np.maximum(np.abs(g["twiss"]), 1e-3))))
    assert np.allclose(tw, g["twiss"], rtol=1e-9, atol=1e-12)
    c = gpu_lib.lattice(corrected_fodo()).calcchrom2()
    print("chrom2", c - g["chrom2"])
    assert np.allclose(c[:2], g["chrom2"][:2], rtol=0, atol=1e-7)
    assert np.allclose(c[2:], g["chrom2"][2:], rtol=1e-4, atol=0)


def test_closed_orbits(gpu_lib, oracle):
    """getco6 against the reference (it inverts I - M by SVD, this by
    elimination; (I - M) has a condition number of a few hundred from the
    slow synchrotron motion, so 1e-10 of the orbit); getco4 against the
    oracle's same computation and by closure (the reference's getco4 returns
    the one-turn image of the off-momentum origin, not a closed orbit)."""
    g = golden("optics2.npz")
    L = gpu_lib.lattice(lat.da_ring(0))
    L.setrfphase()
    co6 = L.getco6()
    assert np.all(np.abs(co6 - g["co6"]) <= 1e-10 * np.abs(g["co6"]).max()), co6 - g["co6"]
    arr = L.get()
    assert all(e.rad == 1 for e in arr if e.type == el.BENDMPOLE)
    assert all(e.on == 1 for e in arr if e.type == el.CAVITY)
    fodo = lat.fodo_ring()
    L = gpu_lib.lattice(fodo)
    for dp in (1e-4, 1e-3):
        co = L.getco4(dp)
        ref = oracle.getco4(el.to_array(fodo), dp)
        # the difference matrix of 1e-8 rays on a 1e-4..1e-3 orbit keeps
        # eps * orbit / 1e-8 ~ 1e-12..1e-11 relative
        assert np.all(np.abs(co - ref) <= 1e-9 * np.abs(ref).max())
        r = np.array([[co[0], co[1], co[2], co[3], 0.0, dp]])
        assert np.abs(L.track(r, 1)[0, :4] - r[0, :4]).max() <= 50 * co[0] ** 2


def test_momentum_aperture_matches_reference(gpu_lib):
    """All locations in one batched launch of the ring rotated to start at
    each; the apertures agree with the reference's except where a particle
    at the edge of the bucket survives on one side of round-off (one step)."""
    g = golden("optics2.npz")
    L = gpu_lib.lattice(lat.da_ring(0))
    L.setrfphase()
    arr = L.get()
    for i in range(len(arr)):
        if arr[i].type == el.BENDMPOLE:
            arr[i].rad = 1
        if arr[i].type == el.CAVITY:
            arr[i].on = 1
    L.update(arr)
    avg, p, m = L.trackma(0.01, 0.0005, 64, 19, g["ma_locs"])
    step = (0.01 - 0.0005) / 19
    print("MA", avg, p, m, "reference", g["ma_avg"], g["ma_plus"], g["ma_minus"])
    assert np.all(np.abs(p - g["ma_plus"]) <= step * (1 + 1e-9))
    assert np.all(np.abs(m - g["ma_minus"]) <= step * (1 + 1e-9))
    if np.array_equal(p, g["ma_plus"]) and np.array_equal(m, g["ma_minus"]):
        assert avg == pytest.approx(g["ma_avg"][0], rel=1e-15)
    with pytest.raises(abi.AccpError, match="location out of range"):
        L.trackma(0.01, 0.0005, 8, 19, [len(arr)])


def test_adjoint_responses_combined_function_bends(gpu_lib):
    """Tune and chromaticity responses by the adjoint on a ring whose bends
    carry a gradient and a sextupole component (PolyB[1], PolyB[2] curvature
    terms in every kick), against fourth-order central differences of the
    GPU's own tunes and chromaticities, as in test_adjoint_responses."""
    ring = lat.combined_function_ring()
    qf, qd, sf, sd = fams(ring, "QF", "QD", "SF", "SD")
    L = gpu_lib.lattice(ring)
    for quantity, f1, f2, field, idx in ((TUNE, qf, qd, el.FIELD_K, 0), (CHROM, sf, sd, el.FIELD_POLYB, 2)):
        adj = L.calcresp(quantity, f1, f2, ADJ)
        measure = L.calctune if quantity == TUNE else L.calcchrom
        for k, fam in enumerate((f1, f2)):
            base = ring[fam[0]].k if field == el.FIELD_K else ring[fam[0]].polyB[idx]
            h = (1e-4 if quantity == TUNE else 1e-2) * abs(base)
            v = {}
            for s in (-2, -1, 1, 2):
                L.update(lat.shift_knob(ring, field, idx, fam, s * h))
                v[s] = measure()
            L.update(ring)
            fd = (-v[2] + 8 * v[1] - 8 * v[-1] + v[-2]) / (12 * h)
            print(quantity, k, "adjoint", adj[k], "fd", fd)
            assert np.all(np.abs(adj[k] - fd) <= 1e-6 * np.abs(fd).max())

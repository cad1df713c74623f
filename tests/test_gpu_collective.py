"""Longitudinal resonator wake and radiation on the GPU.

The wake element has no reference counterpart: it is checked against the
oracle's binned restatement (same histogram and potential, so agreement to
the tracking tolerance), against the brute-force pairwise sum (binning
error only), and the resonator's closed form against its defining
properties.  Radiation with quantum excitation is checked on the
equilibrium of a long run: the GPU and the oracle draw the same
counter-based normal stream, so their equilibrium energy spreads and beam
sizes agree within 2% (Sands' formula sigma_delta^2 = Cq gamma^2 / (J_e rho)
is printed beside them)."""
import math

import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from helpers import assert_coords_close

pytestmark = pytest.mark.gpu


def wake_elem(on=1, nbins=256):
    return el.longwake(5e3, 1.0, 10e9, 2e-9, nbins, -0.015, 0.015, energy=3000.0, on=on)


@pytest.mark.parametrize("nbins", [256, 1024])
def test_wake_kick_matches_oracle_binned(gpu_lib, oracle, nbins):
    """One pass and two passes in a row (the counts of the first must not
    leak into the second) against the oracle; 256 bins form the potential
    inside the kick kernel, 1024 in a kernel of their own."""
    beam = lat.gaussian_beam(20000, seed=8)
    beam[::1000, 0] = np.nan  # lost particles neither kick nor get kicked
    w = [wake_elem(nbins=nbins)]
    L = gpu_lib.lattice(w)
    out = L.track(beam, 1)
    ref = oracle.track(w, beam, 1)
    dd = out[:, 5] - beam[:, 5]
    assert np.abs(np.nan_to_num(dd)).max() > 1e-7  # the kick is there
    assert_coords_close(out, ref, 1e-12, what="wake kick")
    assert_coords_close(L.track(beam, 2), oracle.track(w, beam, 2), 1e-12, what="two wake passes")
    off = gpu_lib.lattice([wake_elem(on=0, nbins=nbins)]).track(beam, 1)
    assert np.array_equal(np.nan_to_num(off), np.nan_to_num(beam))


def test_wake_lattice_tracks_whole_beam(gpu_lib, oracle):
    """A wake inside a ring: the turn is split at the wake so every particle
    reaches it first; 3 turns agree with the oracle."""
    ring = lat.da_ring(0)
    ring.insert(120, wake_elem())
    beam = lat.gaussian_beam(3000, seed=10)
    out = gpu_lib.lattice(ring).track(beam, 3)
    ref = oracle.track(ring, beam, 3)
    assert_coords_close(out, ref, 1e-10, what="ring with wake, 3 turns")


def test_wake_lattice_monitor_and_loss_turns(gpu_lib, oracle):
    """Monitor records and loss turns through a lattice whose turns split
    at a wake, against the oracle turn by turn: statistics of the beam at
    the start of turns 0, 4, 8 (lost particles in their frozen state), the
    turn each particle was lost in (-1 lost on input, 9 survived)."""
    ring = lat.da_ring(0)
    ring.insert(120, wake_elem())
    beam = lat.gaussian_beam(3000, seed=10)
    beam[::250, 0] = np.nan  # lost on input
    beam[7:11, 0] = 0.024  # inside the aperture, lost within a few turns
    beam[7:11, 1] = 2e-3
    L = gpu_lib.lattice(ring)
    out, recs, lt = L.track(beam, 9, mon_interval=4, loss_turn=True)
    r = beam.copy()
    ref_lt = np.where(np.isnan(beam[:, 0]), -1, 9)
    assert len(recs) == 3
    for t in range(9):
        if t % 4 == 0:
            o = oracle.beam_stat(r, t)
            k = t // 4
            assert recs[k][0] == t and recs[k][1] == o[1]
            np.testing.assert_allclose(recs[k][2:], o[2:], rtol=1e-9, atol=1e-18)
        nxt = oracle.track(ring, r, 1, turn0=t)
        ref_lt[np.isnan(nxt[:, 0]) & ~np.isnan(r[:, 0])] = t
        r = nxt
    assert (ref_lt[7:11] < 9).all() and (ref_lt[7:11] >= 0).all()  # the run loses them
    assert np.array_equal(lt, ref_lt), (lt[ref_lt != lt], ref_lt[ref_lt != lt])
    assert_coords_close(out, r, 1e-10, what="ring with wake, 9 turns")
    # a range holding the wake alone: no tracking launch, the loss turns still set
    _, lt_w = L.track(beam, 3, st=120, ed=121, loss_turn=True)
    assert np.array_equal(lt_w, np.where(np.isnan(beam[:, 0]), -1, 3))


def damping_partition(oracle, ring):
    """J_e = 2 + I4/I2 of the isomagnetic ring: periodic dispersion from the
    one-turn matrix, carried to every bend end with the transfer matrices
    (oracle getm66 over element ranges), I4 = sum D/rho^3 * L (trapezoid per
    bend), I2 = sum L/rho^2."""
    m = oracle.getm66(ring)
    eta0 = np.linalg.solve(np.eye(4) - m[:4, :4], m[:4, 5])
    i2 = i4 = 0.0
    for k, e in enumerate(ring):
        if e.type == el.BENDMPOLE:
            def disp(idx):
                if idx == 0:
                    return eta0[0]
                mk = oracle.getm66(ring, ed=idx)
                return (mk[:4, :4] @ eta0 + mk[:4, 5])[0]
            rho = e.len / e.theta
            i2 += e.len / rho ** 2
            i4 += 0.5 * (disp(k) + disp(k + 1)) * e.len / rho ** 3
    return 2.0 + i4 / i2


def test_radiation_equilibrium_energy_spread(gpu_lib, oracle):
    """RADQE on the 20-cell ring at chromaticity +1 with the cavity at the
    synchronous phase, 512 particles: after 10000 turns (2.5 longitudinal
    damping times) the RMS energy spread sampled every 100 turns over 2000
    more turns matches the oracle's run of the same particles and the same
    normal stream (tests/golden/equilibrium.npz) within 2%; the horizontal
    size too.  Sands' linear formula sigma_delta^2 = Cq gamma^2 / (J_e rho),
    with J_e from the ring's own dispersion, holds the energy spread within
    5%: a sampling error of 3% for 512 particles, while a wrong factor in
    the excitation or the damping would move it by tens of percent."""
    from conftest import golden
    from golden.make_equilibrium import BLOCK, NBLOCK, SEED, WARM, equilibrium_ring, initial_beam

    g = golden("equilibrium.npz")
    ring = equilibrium_ring(oracle)
    L = gpu_lib.lattice(ring)
    r = L.track(initial_beam(), WARM, seed=SEED)
    sig, sigx = [], []
    for k in range(NBLOCK):
        r, rec = L.track(r, BLOCK, seed=SEED, turn0=WARM + BLOCK * k, mon_interval=BLOCK)
        sig.append(rec[0][13])
        sigx.append(rec[0][8])
    sd, sx = float(np.mean(sig)), float(np.mean(sigx))
    sd_ref, sx_ref = float(np.mean(g["sigma_delta"])), float(np.mean(g["sigma_x"]))
    gamma = 3000.0 / 0.51099895
    rho = 1.5 / (2 * math.pi / 40)
    je = damping_partition(oracle, lat.fodo_ring())
    sands = math.sqrt(3.8319e-13 * gamma ** 2 / (je * rho))
    print(f"sigma_delta gpu {sd:.4e} oracle {sd_ref:.4e} (Sands linear {sands:.4e}, J_e {je:.4f}); "
          f"sigma_x gpu {sx:.4e} oracle {sx_ref:.4e}")
    assert abs(sd / sd_ref - 1) < 0.02
    assert abs(sx / sx_ref - 1) < 0.02
    assert abs(sd / sands - 1) < 0.05


def test_transwake_matches_reference(gpu_lib):
    """The sorted-scan transverse wake against the reference's pairwise sum
    (one thread, transwake.npz): the kicks agree to 1e-10 of the largest
    (same sum, other order: round-off grows with the chain of exp(c dz)
    factors); everything else is untouched."""
    from conftest import golden
    from tests.golden.make_golden import transwake_case
    g = golden("transwake.npz")
    b, w = transwake_case()
    out = gpu_lib.lattice([w]).track(b, 1)
    live = ~np.isnan(b[:, 0])
    assert np.array_equal(np.isnan(out[:, 0]), ~live)
    for c in (0, 2, 4, 5):
        assert np.array_equal(out[live, c], b[live, c])
    for c in (1, 3):
        kick, ref = out[live, c] - b[live, c], g["pass1"][live, c] - b[live, c]
        assert np.abs(ref).max() > 1e-8
        assert np.abs(kick - ref).max() <= 1e-10 * np.abs(ref).max(), c


def test_transwake_ring_and_length(gpu_lib, oracle):
    """20 turns of a ring with the wake at its end against the oracle; a
    wake with length acts as half drift, kick of the whole beam, half drift;
    switched off it is a drift."""
    beam = lat.gaussian_beam(400, seed=23)
    # strong enough to move px by 2e-4 over the run, below the threshold where
    # it drives the undamped ring unstable within 20 turns (3e6 does)
    w = el.transwake(0.0, 2.0, 2.0e5, 1.5, 5.0, 2.0e5, 2.0, 3.0, on=1)
    ring = lat.fodo_ring() + [w]
    assert_coords_close(gpu_lib.lattice(ring).track(beam, 20), oracle.track(ring, beam, 20), 1e-9,
                        what="ring with a transverse wake, 20 turns")
    wl = el.transwake(0.6, 2.0, 2.0e5, 1.5, 5.0, 2.0e5, 2.0, 3.0, on=1)
    split = [el.drift(0.3), w, el.drift(0.3)]
    # the scanned kick differs from the pairwise sum by round-off (<= 1e-10 of
    # the kick, as against the reference); the second half drift carries it
    # into x
    assert_coords_close(gpu_lib.lattice([wl]).track(beam, 1), oracle.track(split, beam, 1), 1e-10,
                        what="wake with length")
    off = el.transwake(0.6, 2.0, 2.0e5, 1.5, 5.0, 2.0e5, 2.0, 3.0, on=0)
    assert_coords_close(gpu_lib.lattice([off]).track(beam, 1), oracle.track([el.drift(0.6)], beam, 1), 1e-14,
                        what="wake off")


def test_transwake_degenerate_beams(gpu_lib, oracle):
    """TransWake on an all-lost beam (nothing moves), a single particle (no
    other particle, no kick; only the half drifts) and, for a wake of zero
    length, particles sharing one z (sin 0 = 0: no kick between them, bit
    for bit)."""
    w = el.transwake(0.4, 2.0, 1.0e6, 1.5, 5.0, 2.0e6, 2.0, 3.0, on=1)
    L = gpu_lib.lattice([w])
    lost = lat.gaussian_beam(5, seed=3)
    lost[:, 0] = np.nan
    assert np.array_equal(np.isnan(L.track(lost, 3)), np.isnan(lost))
    one = lat.gaussian_beam(1, seed=4)
    assert_coords_close(L.track(one, 2), oracle.track([el.drift(0.4)], one, 2), 1e-14, what="one particle")
    same = lat.gaussian_beam(6, seed=5)
    same[:, 4] = 1e-3
    thin = el.transwake(0.0, 2.0, 1.0e6, 1.5, 5.0, 2.0e6, 2.0, 3.0, on=1)
    assert np.array_equal(gpu_lib.lattice([thin]).track(same, 1), same)


def test_wake_physics_pins_on_gpu(gpu_lib):
    """The GPU's LongWake against the resonator impedance (no oracle in the
    loop): a lone particle sees its half wake q k, a trailer adds the
    leader's W(dz/c), and a Gaussian bunch loses Q^2 k(sigma) within the
    binning error (1e-4)."""
    from test_oracle_math import (check_wake_gaussian_loss_factor, check_wake_single_particle,
                                  check_wake_two_particles)

    def track(elems, r):
        return gpu_lib.lattice(elems).track(r, 1)
    check_wake_single_particle(track)
    check_wake_two_particles(track)
    check_wake_gaussian_loss_factor(track)


def test_collective_workload_against_oracle(gpu_lib, oracle):
    """The collective workload as one piece: the seed-0 DA lattice with
    quantum excitation in every bend (atpass.cpp:255-339 with the 2nd-order
    integrator, accel.cpp:693-700), the cavity at the synchronous phase from
    setrfphase (aline.cpp:465-501) and the LongWake (5 kOhm, Q = 1, 10 GHz,
    2 nC, 256 bins), 32768 particles.  The GPU's setrfphase gives the
    oracle's phase; one turn matches the oracle on the same normal stream at
    1e-12; over 8000 turns (two longitudinal damping times) the energy
    spread, the bunch length and the betatron emittance from the median
    action, averaged over the last 2000 turns, match the oracle's run
    (tests/golden/collective.npz) within 2%, and the survivors within 2% at
    every record.  The lattice's uncorrected chromaticity and 25 mm aperture
    lose particles all along (a quarter of the bunch by turn 8000), so this
    is the loss-shaped state of this lattice, not a clean equilibrium: the
    two runs share their noise but not the last bits of every product, and
    the few halo particles near the aperture whose survival the difference
    decides carry much of the second moments.  An oracle run of 4096
    particles whose initial x moved by one ulp shifted the median-action
    emittance by 1.9% and the energy spread and bunch length by 0.4-0.6%,
    which is why the fixture holds eight times as many particles.  The rms
    emittances (projected and betatron), which the halo carries, are held to
    5%."""
    from conftest import golden
    from golden.make_collective import (EVERY, SEED, TURNS, collective_ring, initial_beam, optics_at_start, record,
                                        wake)

    g = golden("collective.npz")["records"]
    ring = collective_ring(oracle)
    Lp = gpu_lib.lattice(lat.da_ring(0) + [wake()])
    phis = Lp.setrfphase()
    assert abs(phis - ring[-3].phis) <= 1e-12 * abs(phis), (phis, ring[-3].phis)
    L = gpu_lib.lattice(ring)
    beam = initial_beam()
    assert_coords_close(L.track(beam, 1, seed=SEED), oracle.track(ring, beam, 1, seed=SEED), 1e-12,
                        what="collective workload, one turn")
    optics = optics_at_start(oracle, ring)
    r, recs = beam, []
    for k in range(TURNS // EVERY):
        r = L.track(r, EVERY, seed=SEED, turn0=k * EVERY)
        recs.append(record(gpu_lib.beam_stat(r), r, (k + 1) * EVERY, optics))
    recs = np.array(recs)
    assert recs.shape == g.shape and np.array_equal(recs[:, 0], g[:, 0])
    print("turn, survivors gpu/oracle:", [(int(a[0]), int(a[1]), int(b[1])) for a, b in zip(recs[::4], g[::4])])
    assert np.all(np.abs(recs[:, 1] / g[:, 1] - 1) <= 0.02)
    last = slice(-(2000 // EVERY), None)
    for c, name, tol in ((2, "sigma_delta", 0.02), (3, "sigma_z", 0.02), (8, "emittance_x (median action)", 0.02),
                         (7, "emittance_x (betatron rms)", 0.05), (4, "emittance_x (projected rms)", 0.05)):
        a, b = recs[last, c].mean(), g[last, c].mean()
        print(f"{name}: gpu {a:.5e} oracle {b:.5e} ({a / b - 1:+.2%}, bound {tol:.0%})")
        assert abs(a / b - 1) <= tol, name

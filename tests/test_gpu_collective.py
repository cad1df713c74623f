"""Longitudinal resonator wake and radiation on the GPU.

The wake element has no reference counterpart: it is checked against the
oracle's binned restatement (same histogram and potential, so agreement to
the tracking tolerance), against the brute-force pairwise sum (binning
error only), and the resonator's closed form against its defining
properties.  Radiation with quantum excitation is checked statistically:
the equilibrium energy spread of a long run against Sands' isomagnetic
formula sigma_delta^2 = Cq gamma^2 / (J_e rho)."""
import math

import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from helpers import assert_coords_close

pytestmark = pytest.mark.gpu


def wake_elem(on=1):
    return el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.015, 0.015, energy=3000.0, on=on)


def test_wake_kick_matches_oracle_binned(gpu_lib, oracle):
    beam = lat.gaussian_beam(20000, seed=8)
    beam[::1000, 0] = np.nan  # lost particles neither kick nor get kicked
    w = [wake_elem()]
    out = gpu_lib.lattice(w).track(beam, 1)
    ref = oracle.track(w, beam, 1)
    dd = out[:, 5] - beam[:, 5]
    assert np.abs(np.nan_to_num(dd)).max() > 1e-7  # the kick is there
    assert_coords_close(out, ref, 1e-12, what="wake kick")
    off = gpu_lib.lattice([wake_elem(on=0)]).track(beam, 1)
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
    """RADQE on the 20-cell ring with the cavity at the synchronous phase:
    after 10000 turns (2.5 longitudinal damping times) the energy spread,
    averaged over 2000 more turns, is Sands' isomagnetic value
    sigma_delta^2 = Cq gamma^2 / (J_e rho) within 2%."""
    ring = lat.da_ring(0)
    arr = el.to_array(ring)
    oracle.setrfphase(arr)
    ring = [arr[i] for i in range(len(arr))]
    lat.set_rad(ring, 2)
    lat.set_cavity(ring, 1)
    beam = lat.gaussian_beam(4096, seed=12) * np.array([1, 1, 1, 1, 1, 0.5])
    L = gpu_lib.lattice(ring)
    r = L.track(beam, 10000, seed=77)
    sig = []
    for k in range(20):
        r, rec = L.track(r, 100, seed=77, turn0=10000 + 100 * k, mon_interval=100)
        sig.append(rec[0][13])
    sd = float(np.mean(sig))
    gamma = 3000.0 / 0.51099895
    rho = 1.5 / (2 * math.pi / 40)
    cq = 3.8319e-13
    bare = [e for e in lat.fodo_ring()]
    je = damping_partition(oracle, bare)
    sands = math.sqrt(cq * gamma ** 2 / (je * rho))
    print(f"equilibrium sigma_delta {sd:.4e}, Sands {sands:.4e} (J_e = {je:.4f})")
    assert abs(sd / sands - 1) < 0.02

"""The collective workload on the oracle: the seed-0 DA lattice (cavity,
aperture, multipole errors) with quantum excitation in every bend, the
cavity at the synchronous phase from setrfphase, and the longitudinal
resonator wake (5 kOhm, Q = 1, 10 GHz, 2 nC in 256 bins over +-15 cm) at the
end of the ring; 32768 Gaussian particles, 8000 turns (two longitudinal
damping times), with the counter-based normal stream the GPU draws.  Every
250 turns it records the survivors, the energy spread, the bunch length, the
projected horizontal emittance, the bunch centroid, and the betatron
emittance (dispersion taken out) twice: from the second moments and from
the median of the betatron action.  Writes tests/golden/collective.npz.

The wake couples the particles, so each turn runs the elements before the
wake in parallel over particles (they are independent there) and then the
wake over the whole beam; the wake is the last element, so every other
element keeps the index its random numbers are keyed on.  Takes about 70
minutes on 8 cores: python tests/golden/make_collective.py"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "accel-plus_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyaccp.elements as el  # noqa: E402
import pyaccp.lattices as lat  # noqa: E402
from pyoracle import Oracle  # noqa: E402

N, SEED, TURNS, EVERY = 32768, 3, 8000, 250


def wake():
    return el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.15, 0.15, energy=3000.0, on=1)


def collective_ring(oracle):
    """The DA lattice with the wake last, phase from the oracle's
    setrfphase (the wake sits out its probe turn), RADQE and the cavity on."""
    arr = el.to_array(lat.da_ring(0) + [wake()])
    oracle.setrfphase(arr)
    ring = [arr[i] for i in range(len(arr))]
    lat.set_rad(ring, 2)
    lat.set_cavity(ring, 1)
    assert ring[-1].type == el.LONGWAKE and ring[-1].on == 1
    return ring


def initial_beam():
    return lat.gaussian_beam(N, seed=13)


def optics_at_start(oracle, ring):
    """beta_x, alpha_x and the dispersion (eta_x, eta_x') at the ring start
    from the oracle's one-turn matrix of the ring without radiation, cavity
    and wake."""
    lin = [e.copy() for e in ring]
    for e in lin:
        if e.type == el.BENDMPOLE:
            e.rad = 0
        if e.type in (el.CAVITY, el.LONGWAKE):
            e.on = 0
    m = oracle.getm66(lin)
    eta = np.linalg.solve(np.eye(4) - m[:4, :4], m[:4, 5])
    c = 0.5 * (m[0, 0] + m[1, 1])
    s = np.copysign(np.sqrt(1.0 - c * c), m[0, 1])
    return m[0, 1] / s, (m[0, 0] - m[1, 1]) / (2.0 * s), eta[0], eta[1]


def record(stat, r, turn, optics):
    """turn, survivors, sigma_delta, sigma_z, projected horizontal emittance,
    <z>, <delta> from a calc_beam_stat record; then the betatron emittance
    of the live particles (x - eta delta, x' - eta' delta) from the second
    moments, and from the median betatron action: for a Gaussian beam the
    action J is exponential with mean eps, so eps = median(J) / ln 2 (a
    handful of halo particles near the aperture barely moves it)."""
    emit = np.sqrt(max(stat[8] ** 2 * stat[9] ** 2 - stat[14] ** 2, 0.0))
    beta, alpha, eta, etap = optics
    live = r[~np.isnan(r[:, 0])]
    d = live[:, 5] - live[:, 5].mean()
    xb = live[:, 0] - live[:, 0].mean() - eta * d
    pb = live[:, 1] - live[:, 1].mean() - etap * d
    e_rms = np.sqrt(max(np.mean(xb * xb) * np.mean(pb * pb) - np.mean(xb * pb) ** 2, 0.0))
    gamma = (1.0 + alpha * alpha) / beta
    action = 0.5 * (gamma * xb * xb + 2.0 * alpha * xb * pb + beta * pb * pb)
    return [turn, stat[1], stat[13], stat[12], emit, stat[6], stat[7], e_rms, np.median(action) / np.log(2.0)]


def main():
    O = Oracle()
    ring = collective_ring(O)
    optics = optics_at_start(O, ring)
    body, wk = ring[:-1], [ring[-1]]
    r = initial_beam()
    recs = []
    t0 = time.time()
    threads = os.cpu_count() or 8
    for t in range(TURNS):
        r = O.track_omp(body, r, 1, threads, seed=SEED, turn0=t)
        r = O.track(wk, r, 1)
        if (t + 1) % EVERY == 0:
            recs.append(record(O.beam_stat(r), r, t + 1, optics))
            print(f"turn {t + 1}: {recs[-1][1]:.0f} left, sigma_delta {recs[-1][2]:.4e}, "
                  f"sigma_z {recs[-1][3]:.4e}, emittance {recs[-1][4]:.4e}, betatron {recs[-1][7]:.4e} "
                  f"(median action {recs[-1][8]:.4e}) ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(os.path.join(HERE, "collective.npz"), records=np.array(recs), n=np.array([N]),
                        seed=np.array([SEED]), turns=np.array([TURNS]), every=np.array([EVERY]))


if __name__ == "__main__":
    main()

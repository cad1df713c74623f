"""The collective workload on the oracle: the seed-0 DA lattice (cavity,
aperture, multipole errors) with quantum excitation in every bend, the
cavity at the synchronous phase from setrfphase, and the longitudinal
resonator wake (5 kOhm, Q = 1, 10 GHz, 2 nC in 256 bins over +-15 cm) at the
end of the ring; 4096 Gaussian particles, 8000 turns (two longitudinal
damping times), with the counter-based normal stream the GPU draws.  Every
250 turns it records the survivors, the energy spread, the bunch length, the
horizontal emittance and the bunch centroid.  Writes
tests/golden/collective.npz.

The wake couples the particles, so each turn runs the elements before the
wake in parallel over particles (they are independent there) and then the
wake over the whole beam; the wake is the last element, so every other
element keeps the index its random numbers are keyed on.  Takes about ten
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

N, SEED, TURNS, EVERY = 4096, 3, 8000, 250


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


def record(stat, turn):
    """turn, survivors, sigma_delta, sigma_z, horizontal emittance, <z>, <delta>
    from a calc_beam_stat record."""
    emit = np.sqrt(max(stat[8] ** 2 * stat[9] ** 2 - stat[14] ** 2, 0.0))
    return [turn, stat[1], stat[13], stat[12], emit, stat[6], stat[7]]


def main():
    O = Oracle()
    ring = collective_ring(O)
    body, wk = ring[:-1], [ring[-1]]
    r = initial_beam()
    recs = []
    t0 = time.time()
    threads = os.cpu_count() or 8
    for t in range(TURNS):
        r = O.track_omp(body, r, 1, threads, seed=SEED, turn0=t)
        r = O.track(wk, r, 1)
        if (t + 1) % EVERY == 0:
            recs.append(record(O.beam_stat(r), t + 1))
            print(f"turn {t + 1}: {recs[-1][1]:.0f} left, sigma_delta {recs[-1][2]:.4e}, "
                  f"sigma_z {recs[-1][3]:.4e}, emittance {recs[-1][4]:.4e} ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(os.path.join(HERE, "collective.npz"), records=np.array(recs), n=np.array([N]),
                        seed=np.array([SEED]), turns=np.array([TURNS]), every=np.array([EVERY]))


if __name__ == "__main__":
    main()

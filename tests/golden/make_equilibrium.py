"""Radiation equilibrium of the oracle (the reference's arithmetic, with the
counter-based normal stream the GPU uses): RADQE, cavity on at the
synchronous phase, 512 particles, 10000 turns, then the RMS energy spread
every 100 turns for 2000 more.  Writes tests/golden/equilibrium.npz.
Takes a few minutes on 8 cores: python tests/golden/make_equilibrium.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "accel-plus_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyaccp.elements as el  # noqa: E402
import pyaccp.lattices as lat  # noqa: E402
from pyoracle import Oracle  # noqa: E402

N, SEED, WARM, BLOCK, NBLOCK = 512, 77, 10000, 100, 20


# sextupoles for chromaticity +1 in both planes (the oracle's calcchrom,
# linear in the two strengths); the ring's own SF = 20, SD = -25 give +101 and
# +54, which carry off-momentum particles across resonances within a few
# thousand turns
SF, SD = 1.34541125, -2.13767167


def equilibrium_ring(oracle):
    """The 20-cell FODO ring at chromaticity +1, the 1.5 MV cavity at the
    synchronous phase (the oracle's setrfphase), the 25 x 12 mm aperture, and
    quantum excitation in every bend: no particle is lost, and tracks that
    share their noise stay together over the whole run."""
    ring = lat.fodo_ring(sf=SF, sd=SD)
    ring.append(el.cavity(0.0, 1.5, lat.RF_FREQ, 0.0, el.DEFAULT_E0, on=1, name="RF"))
    ring.append(el.aperture(0.0, (-0.025, 0.025), (-0.012, 0.012), name="AP"))
    arr = el.to_array(ring)
    oracle.setrfphase(arr)
    ring = [arr[i] for i in range(len(arr))]
    lat.set_rad(ring, 2)
    lat.set_cavity(ring, 1)
    return ring


def initial_beam():
    return lat.gaussian_beam(N, seed=12) * np.array([1, 1, 1, 1, 1, 0.5])


def main():
    O = Oracle()
    ring = equilibrium_ring(O)
    r = O.track_omp(ring, initial_beam(), WARM, os.cpu_count() or 8, seed=SEED)
    sig, sigx = [], []
    for k in range(NBLOCK):
        s = O.beam_stat(r)
        sig.append(s[13])
        sigx.append(s[8])
        r = O.track_omp(ring, r, BLOCK, os.cpu_count() or 8, seed=SEED, turn0=WARM + BLOCK * k)
    np.savez_compressed(os.path.join(HERE, "equilibrium.npz"), sigma_delta=np.array(sig),
                        sigma_x=np.array(sigx), n=np.array([N]), seed=np.array([SEED]),
                        warm=np.array([WARM]), block=np.array([BLOCK]))
    print("sigma_delta", np.mean(sig), "sigma_x", np.mean(sigx), "lost", int(np.isnan(r[:, 0]).sum()))


if __name__ == "__main__":
    main()

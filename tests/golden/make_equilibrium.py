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


def equilibrium_ring(oracle):
    arr = el.to_array(lat.da_ring(0))
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

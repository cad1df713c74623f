"""Radiation equilibrium of the compiled reference against the oracle.

The reference draws its quantum-excitation noise from rand() (grandn,
utility.cpp:194-217), the oracle from the counter-based Philox stream, so
the two runs agree only statistically.  Both track the same 16384-particle
bunch through the RADQE ring with the cavity at the synchronous phase for
10000 turns, then record the beam every 100 turns for 2000 more.  The
reference runs on one thread (its rand() is shared between threads).
Writes tests/golden/equilibrium_ref.npz.  Takes about half an hour:
python tests/golden/make_equilibrium_ref.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "accel-plus_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import pyaccp.lattices as lat  # noqa: E402
from make_equilibrium import equilibrium_ring  # noqa: E402
from pyoracle import Oracle, Reference  # noqa: E402

N, SEED, WARM, BLOCK, NBLOCK = 16384, 5, 10000, 100, 20


def stats(s):
    """sigma_delta, sigma_x and the projected rms emittance
    sqrt(<x x><x' x'> - <x x'>^2) of one beam_stat record."""
    return s[13], s[8], np.sqrt(max(s[8] ** 2 * s[9] ** 2 - s[14] ** 2, 0.0))


def run(step, r):
    r = step(r, WARM, 0)
    rows = []
    for k in range(NBLOCK):
        rows.append(stats(Oracle().beam_stat(r)))
        r = step(r, BLOCK, WARM + BLOCK * k)
    return np.array(rows), int(np.isnan(r[:, 0]).sum())


def main():
    O = Oracle()
    ring = equilibrium_ring(O)
    beam = lat.gaussian_beam(N, seed=21) * np.array([1, 1, 1, 1, 1, 0.5])
    threads = os.cpu_count() or 8
    orc, orc_lost = run(lambda r, n, t0: O.track_omp(ring, r, n, threads, seed=SEED, turn0=t0), beam)
    print("oracle", orc.mean(axis=0), "lost", orc_lost, flush=True)
    line = Reference().load(ring)
    line.set_threads(1)
    ref, ref_lost = run(lambda r, n, t0: line.track(r, n), beam)
    print("reference", ref.mean(axis=0), "lost", ref_lost, flush=True)
    np.savez_compressed(os.path.join(HERE, "equilibrium_ref.npz"), oracle=orc, reference=ref,
                        n=np.array([N]), warm=np.array([WARM]), block=np.array([BLOCK]))


if __name__ == "__main__":
    main()

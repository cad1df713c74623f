"""Reverse-mode gradient of the turn map on the GPU against central finite
differences: of the GPU's own forward map and of the oracle's.  Tolerance:
1e-6 relative per knob.  The adjoint is exact up to round-off; the central
difference carries O(h^2) truncation and a round-off error that grows as
1/h (for the small-gradient knobs a step of 1e-7 already costs ~1e-5), so
the step is 1e-5 of the knob (at least 1e-6): truncation and round-off then
both stay near 1e-8."""
import numpy as np
import pytest

import pyaccp.abi as abi
import pyaccp.elements as el
import pyaccp.lattices as lat

pytestmark = pytest.mark.gpu


def fd_step(ring, f, i, m):
    base = ring[m[0]].k if f == el.FIELD_K else ring[m[0]].polyB[i]
    return max(1e-5 * abs(base), 1e-6)


def loss_and_seed(r):
    return float(np.sum(r * r)), 2.0 * r


def adjoint_lattice():
    ring = lat.da_ring(0, extra_quads=True)
    lat.set_rad(ring, 0)
    lat.set_cavity(ring, 1)
    return ring


def test_gradient_matches_gpu_finite_differences(gpu_lib):
    ring = adjoint_lattice()
    knobs = lat.adjoint_knobs(ring)
    assert len(knobs) == 8
    beam = lat.gaussian_beam(256, seed=1)
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    nturn = 40
    out = A.forward(beam, nturn, ckpt_every=16)
    assert np.array_equal(out, L.track(beam, nturn))  # forward is the tracking map
    loss, seed = loss_and_seed(out)
    grad, dr0 = A.backward(seed)
    for k, (name, f, i, m) in enumerate(knobs):
        h = fd_step(ring, f, i, m)
        lp = gpu_lib.lattice(lat.shift_knob(ring, f, i, m, h))
        lm = gpu_lib.lattice(lat.shift_knob(ring, f, i, m, -h))
        fd = (loss_and_seed(lp.track(beam, nturn))[0] - loss_and_seed(lm.track(beam, nturn))[0]) / (2 * h)
        rel = abs(grad[k] - fd) / max(abs(fd), 1e-30)
        print(f"{name}: adjoint {grad[k]:+.12e}  fd {fd:+.12e}  rel {rel:.1e}")
        assert rel < 1e-6, name
    # dL/dr0 on a few coordinates
    for (pi, c) in ((0, 0), (17, 1), (100, 2), (200, 5)):
        h = 1e-5 * max(abs(beam[pi, c]), 1e-6)
        bp, bm = beam.copy(), beam.copy()
        bp[pi, c] += h
        bm[pi, c] -= h
        fd = (loss_and_seed(L.track(bp, nturn))[0] - loss_and_seed(L.track(bm, nturn))[0]) / (2 * h)
        assert abs(dr0[pi, c] - fd) <= 1e-6 * max(abs(fd), 1e-12), (pi, c, dr0[pi, c], fd)


def test_gradient_matches_oracle_finite_differences(gpu_lib, oracle):
    """Central differences of the oracle (the reference's arithmetic) on a
    small case: 16 particles, 12 turns, checkpoints every 5 turns."""
    ring = adjoint_lattice()
    knobs = lat.adjoint_knobs(ring)
    beam = lat.gaussian_beam(16, seed=3)
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    out = A.forward(beam, 12, ckpt_every=5)
    _, seed = loss_and_seed(out)
    grad, _ = A.backward(seed)
    for k, (name, f, i, m) in enumerate(knobs):
        h = fd_step(ring, f, i, m)
        fp = loss_and_seed(oracle.track(lat.shift_knob(ring, f, i, m, h), beam, 12))[0]
        fm = loss_and_seed(oracle.track(lat.shift_knob(ring, f, i, m, -h), beam, 12))[0]
        fd = (fp - fm) / (2 * h)
        assert abs(grad[k] - fd) <= 1e-6 * abs(fd), (name, grad[k], fd)


def test_adjoint_rejects_unsupported_lattices(gpu_lib):
    ring = lat.fodo_ring()
    lat.set_rad(ring, 1)
    L = gpu_lib.lattice(ring)
    with pytest.raises(abi.AccpError, match="no reverse map"):
        abi.Adjoint(gpu_lib, L, [])
    L2 = gpu_lib.lattice(lat.fodo_ring())
    A = abi.Adjoint(gpu_lib, L2, [(el.FIELD_K, 0, lat.family(lat.fodo_ring(), "QF"))])
    with pytest.raises(abi.AccpError, match="before accp_adjoint_forward"):
        A.backward(np.zeros((4, 6)))

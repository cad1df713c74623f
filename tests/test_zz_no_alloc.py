"""accp_track_device_noalloc (no device allocation; the working set lives in
a fixed library array) against accp_track_device on the same inputs:
coordinates and loss turns equal bit for bit, on two calls in a row, the
second with a smaller beam, and on a ring whose turn splits at a LongWake.
Beyond its limit, and with a TransWake, it refuses with an error before
touching device memory."""
import numpy as np
import pytest

import pyaccp.abi as abi
import pyaccp.elements as el
import pyaccp.lattices as lat
from helpers import same_bits

pytestmark = pytest.mark.gpu


def run(L, beam, nturn, noalloc, **kw):
    from pyaccp.hipmem import DeviceArray, synchronize
    d = DeviceArray(beam)
    d_lt = DeviceArray(np.zeros(len(beam), dtype=np.int32))
    L.track_device(d.ptr, len(beam), nturn, d_loss_turn=d_lt.ptr, noalloc=noalloc, **kw)
    synchronize()
    return d.numpy(), d_lt.numpy()


def test_noalloc_equals_allocating_entry_point(gpu_lib):
    L = gpu_lib.lattice(lat.da_ring(0))
    grid = lat.da_grid(41, 21)  # 4305 points: losses and compaction between chunks
    small = lat.gaussian_beam(300, seed=4)
    for beam, nturn in ((grid, 96), (small, 20)):
        a, la = run(L, beam, nturn, False)
        b, lb = run(L, beam, nturn, True)
        print(f"{len(beam)} particles, {nturn} turns: {int((~np.isnan(b[:, 0])).sum())} survive")
        assert same_bits(a, b) and np.array_equal(la, lb)
    # element range and turn offset, a launch shape of its own
    t = abi.Tuning(128, 2, 2, 8)
    a, _ = run(L, small, 7, False, st=5, ed=200, turn0=3, tuning=t)
    b, _ = run(L, small, 7, True, st=5, ed=200, turn0=3, tuning=t)
    assert same_bits(a, b)
    # a LongWake: the turn splits at the wake, whose buffers the lattice holds
    ring = lat.da_ring(0) + [el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.15, 0.15, energy=3000.0, on=1)]
    lat.set_rad(ring, 2)
    Lw = gpu_lib.lattice(ring)
    beam = lat.gaussian_beam(2000, seed=6)
    a, la = run(Lw, beam, 5, False, seed=11)
    b, lb = run(Lw, beam, 5, True, seed=11)
    assert same_bits(a, b) and np.array_equal(la, lb)


def test_noalloc_limits(gpu_lib):
    from pyaccp.hipmem import DeviceArray
    L = gpu_lib.lattice(lat.da_ring(0))
    d = DeviceArray(lat.gaussian_beam(4, seed=1))
    with pytest.raises(abi.AccpError, match="at most"):
        L.track_device(d.ptr, (1 << 20) + 1, 1, noalloc=True)
    tw = lat.fodo_ring() + [el.transwake(0.0, 2.0, 2.0e5, 1.5, 5.0, 2.0e5, 2.0, 3.0, on=1)]
    with pytest.raises(abi.AccpError, match="TransWake"):
        gpu_lib.lattice(tw).track_device(d.ptr, 4, 1, noalloc=True)

"""The lattice-specialized tracker (ACCP_STAGE_SPECIALIZED): the turn's group
sequence compiled at run time into a kernel of its own.

Its element maps are the table walker's, in the same order.  Written out as
straight-line code, a product at the end of one map may be fused with a sum
at the start of the next (the compiler's FMA contraction, which the walker's
per-group branches keep apart), so the two differ in the last bits: the
specialized kernel is held to the oracle at the tracking tolerance (1e-12 per
coordinate after one turn, ALine::pass, aline.cpp:359-363), to the oracle's
DA survival maps except next to the boundary (trackDA's rule,
aline.cpp:607-715), and its batched, captured and plain runs to each other
bit for bit."""
import numpy as np
import pytest

import pyaccp.abi as abi
import pyaccp.elements as el
import pyaccp.lattices as lat
from helpers import assert_coords_close, same_bits

pytestmark = pytest.mark.gpu

SPEC = abi.STAGE_SPECIALIZED


def spec(block=64, ppl=1, compact=0):
    return abi.Tuning(block, ppl, SPEC, compact)


@pytest.mark.parametrize("block,ppl", [(64, 1), (256, 1), (128, 2)])
def test_specialized_matches_oracle(gpu_lib, oracle, block, ppl):
    ring = lat.da_ring(1, strength_errors=True)
    beam = lat.gaussian_beam(777, seed=3)
    L = gpu_lib.lattice(ring)
    print(f"block {block}, ppl {ppl}")
    err = assert_coords_close(L.track(beam, 1, tuning=spec(block, ppl)), oracle.track(ring, beam, 1), 1e-12,
                              what=f"{block}/{ppl} one turn")
    print("one turn, max scaled difference", err)
    assert_coords_close(L.track(beam, 2, tuning=spec(block, ppl)), oracle.track(ring, beam, 2), 1e-11,
                        what=f"{block}/{ppl} two turns")


def test_specialized_rings(gpu_lib, oracle):
    """The FODO ring (20 cells, no cavity or aperture), the adjoint
    workload's ring with its extra quad families (cells of two kinds: the
    repeated block covers part of the ring) and a short line with no
    repetition at all: one turn at 1e-12, ten at 1e-10."""
    beam = lat.gaussian_beam(1000, seed=0)
    short = lat.da_ring(2)[:23] + [el.aperture(0.0, (-0.025, 0.025), (-0.012, 0.012))]
    for name, ring in (("fodo", lat.fodo_ring()), ("da + extra quads", lat.da_ring(0, extra_quads=True)),
                       ("short line", short)):
        L = gpu_lib.lattice(ring)
        print(name, len(ring), "elements")
        assert_coords_close(L.track(beam, 1, tuning=spec()), oracle.track(ring, beam, 1), 1e-12, what=name)
        g, o = L.track(beam, 10, tuning=spec()), oracle.track(ring, beam, 10)
        fg, fo = np.isfinite(g).all(axis=1), np.isfinite(o).all(axis=1)
        assert np.array_equal(fg, fo)
        assert_coords_close(g[fg], o[fo], 1e-10, what=f"{name}, 10 turns")


def test_specialized_da_survival_map(gpu_lib, oracle):
    """Seed-0 DA lattice, 21 x 11 x 5 grid, 256 turns, with compaction
    between chunks and loss turns recorded: survival identical to the
    oracle's except next to the boundary; the loss turns of particles lost
    away from the boundary are the walker's."""
    da = lat.da_ring(0)
    grid = lat.da_grid(21, 11)
    L = gpu_lib.lattice(da)
    g, lt = L.track(grid, 256, tuning=spec(), loss_turn=True)
    w, lw = L.track(grid, 256, loss_turn=True)
    o = oracle.track(da, grid, 256)
    sg = ~np.isnan(g[:, 0]).reshape(5, 11, 21)
    so = ~np.isnan(o[:, 0]).reshape(5, 11, 21)
    for d, iy, ix in np.argwhere(sg != so):
        nb = so[d, max(iy - 1, 0):iy + 2, max(ix - 1, 0):ix + 2]
        assert nb.any() and not nb.all(), f"survival differs away from the boundary at {(d, iy, ix)}"
    live = so.ravel() & sg.ravel()
    assert_coords_close(g[live], o[live], 1e-6, what="survivors after 256 turns")
    same = (lt == lw)
    print("loss turns equal to the walker's:", int(same.sum()), "of", len(lt))
    assert (~same).sum() <= 5


def test_specialized_batched_plain_and_graph_agree(gpu_lib):
    """Four error seeds with uneven beams (one of a single particle): the
    batched launch equals one plain call per seed bit for bit, and replays
    of the captured graph, with eager work between them, equal the eager
    batched run."""
    from pyaccp.hipmem import DeviceArray, synchronize

    seeds = [lat.da_ring(s, strength_errors=True) for s in range(4)]
    sizes = [130, 64, 1, 257]
    beams = [lat.gaussian_beam(n, seed=20 + i) * 20 for i, n in enumerate(sizes)]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    lats = [gpu_lib.lattice(s) for s in seeds]
    t = spec(64, 1, 8)
    plain = np.concatenate([L.track(b, 40, tuning=t) for L, b in zip(lats, beams)])
    B = abi.Batch(gpu_lib, lats)
    beam = np.concatenate(beams)
    d = DeviceArray(beam)
    B.track(d.ptr, offs, 40, tuning=t)
    synchronize()
    batched = d.numpy()
    print("survivors after 40 turns:", int((~np.isnan(batched[:, 0])).sum()), "of", len(beam))
    assert same_bits(batched, plain)
    dg = DeviceArray(beam)
    B.capture(dg.ptr, offs, 40, tuning=t)
    other = gpu_lib.lattice(lat.fodo_ring())
    for rep in range(3):
        dg.upload(beam)
        synchronize()
        B.replay()
        synchronize()
        other.track(lat.gaussian_beam(50, seed=rep), 1)
        assert same_bits(dg.numpy(), batched), f"replay {rep}"


def test_specialized_monitor_records(gpu_lib):
    """Monitor records (turn-start snapshots) from the specialized kernel
    agree with the walker's to round-off."""
    ring = lat.da_ring(0)
    beam = lat.gaussian_beam(2000, seed=6)
    beam[::97, 0] = np.nan
    L = gpu_lib.lattice(ring)
    _, a = L.track(beam, 9, mon_interval=4, tuning=spec())
    _, b = L.track(beam, 9, mon_interval=4)
    assert len(a) == len(b) == 3
    np.testing.assert_array_equal(a[:, :2], b[:, :2])
    np.testing.assert_allclose(a[:, 2:], b[:, 2:], rtol=1e-11, atol=1e-20)


def test_specialized_full_variant_and_ranges(gpu_lib, oracle):
    """Lattices of the full tracker variant and element ranges: radiation
    (flag 1) and quantum excitation on the oracle's random stream, one turn
    at 1e-12; an element range and a turn offset at 1e-12; a ring whose turn
    splits at a LongWake (one kernel per segment), 3 turns at 1e-10; the
    element types beyond the tracking set (Matrix, SyncRad, CrabCav, RFMult,
    both wigglers) against the reference's fixtures at 1e-12, and a line of
    the noisy ones against the oracle on the same stream, 3 turns at 1e-12."""
    beam = lat.gaussian_beam(500, seed=2)
    ring = lat.fodo_ring()
    lat.set_rad(ring, 1)
    assert_coords_close(gpu_lib.lattice(ring).track(beam, 1, tuning=spec(256)), oracle.track(ring, beam, 1), 1e-12,
                        what="rad flag 1")
    qe = lat.da_ring(0)
    lat.set_rad(qe, 2)
    kw = dict(seed=1234, turn0=7, pid0=100)
    assert_coords_close(gpu_lib.lattice(qe).track(beam, 1, tuning=spec(), **kw), oracle.track(qe, beam, 1, **kw),
                        1e-12, what="rad flag 2")
    da = lat.da_ring(0, extra_quads=True)
    L = gpu_lib.lattice(da)
    for st, ed in ((0, 1), (17, 100), (100, len(da))):
        assert_coords_close(L.track(beam, 1, st, ed, turn0=3, tuning=spec()),
                            oracle.track(da, beam, 1, st, ed, turn0=3), 1e-12, what=f"range {st}..{ed}")
    wring = lat.da_ring(0)
    wring.insert(120, el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.015, 0.015, energy=3000.0, on=1))
    wbeam = lat.gaussian_beam(3000, seed=10)
    assert_coords_close(gpu_lib.lattice(wring).track(wbeam, 3, tuning=spec()), oracle.track(wring, wbeam, 3), 1e-10,
                        what="ring with a wake, 3 turns")
    from conftest import golden
    from tests.golden.make_golden import elements2_cases
    g = golden("elements2.npz")
    P = g["particles"]
    for name, e in elements2_cases():
        assert_coords_close(gpu_lib.lattice([e]).track(P, 1, tuning=spec()), g[name], 1e-12, what=name)
    noisy = [el.rfmult(0.1, 5e8, 0.2, [1e-4, 2e-3, 0.5], [0.0, 1e-3, 0.2], sigphi=2e-3),
             el.syncrad(0.0, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0, rad=el.RAD_LUMP_QE),
             el.crabcav(0.3, 1.5, 0.5, 1.5e9, 0.3, 3000.0, sigphi=1e-3, sigdvv=1e-2)]
    kw = dict(seed=17, turn0=5, pid0=100)
    assert_coords_close(gpu_lib.lattice(noisy).track(beam, 3, tuning=spec(), **kw),
                        oracle.track(noisy, beam, 3, **kw), 1e-12, what="noisy line, 3 turns")


def test_lattice_default_tuning(gpu_lib, oracle):
    """accp_lattice_set_tuning: calls that name no launch shape take the
    lattice's.  With the specialized kernel as the lattice's default, plain
    tracking equals an explicit specialized call bit for bit; trackDA (which
    tracks internally) keeps the oracle's survival of every ray point except
    next to the boundary; the adjoint, whose forward and recomputation
    passes then run on the specialized kernel, gives the walker's gradient
    to round-off in both reverse forms; resetting
    the default restores the walker's results bit for bit."""
    da = lat.da_ring(0)
    L = gpu_lib.lattice(da)
    beam = lat.gaussian_beam(500, seed=3)
    walker = L.track(beam, 5)
    L.set_tuning(spec())
    assert same_bits(L.track(beam, 5), L.track(beam, 5, tuning=spec()))
    area_s, rx_s, ry_s, lt_s = L.trackda(0.02, 0.001, 0.01, 64, 11, 10, loss_turn=True)
    L.set_tuning(None)
    assert same_bits(L.track(beam, 5), walker)
    area_w, rx_w, ry_w, lt_w = L.trackda(0.02, 0.001, 0.01, 64, 11, 10, loss_turn=True)
    print("trackDA area: specialized", area_s, "walker", area_w, "loss turns differ at", int((lt_s != lt_w).sum()))
    assert abs(area_s - area_w) <= 0.05 * area_w and (lt_s != lt_w).sum() <= 3

    ring = lat.da_ring(0, extra_quads=True)
    lat.set_rad(ring, 0)
    lat.set_cavity(ring, 1)
    knobs = [(f, i, m) for _, f, i, m in lat.adjoint_knobs(ring)]
    b4 = lat.gaussian_beam(512, seed=1) * np.array([0.5, 0.5, 0.5, 0.5, 0.25, 0.25])
    La = gpu_lib.lattice(ring)
    for mode in (abi.Adjoint.REVERSE_SEQUENTIAL, abi.Adjoint.REVERSE_SEGMENT_JACOBIAN):
        grads = []
        for t in (None, spec()):
            La.set_tuning(t)
            A = abi.Adjoint(gpu_lib, La, knobs)
            A.set_reverse(mode)
            out = A.forward(b4, 64, ckpt_every=16)
            g, dr0 = A.backward(np.where(np.isnan(out[:, :1]), 0.0, 2.0 * np.nan_to_num(out)))
            grads.append((g, dr0))
            A.free()
        (gw, dw), (gs, ds) = grads
        print(f"adjoint (reverse mode {mode}), specialized forward vs the walker's:",
              np.max(np.abs(gs - gw) / np.abs(gw)))
        np.testing.assert_allclose(gs, gw, rtol=1e-9)
        np.testing.assert_allclose(ds, dw, rtol=1e-9, atol=1e-12 * np.abs(dw).max())


def test_specialized_refusals(gpu_lib):
    """Calls the specialized kernel does not take fail with an error and
    change nothing: the no-allocation entry point (it would load a code
    object) and a batch whose seeds differ in structure."""
    from pyaccp.hipmem import DeviceArray

    beam = lat.gaussian_beam(64, seed=1)
    L = gpu_lib.lattice(lat.da_ring(0))
    d = DeviceArray(beam)
    with pytest.raises(abi.AccpError, match="noalloc"):
        L.track_device(d.ptr, len(beam), 1, tuning=spec(), noalloc=True)
    assert np.array_equal(d.numpy(), beam)
    mixed = [gpu_lib.lattice(lat.da_ring(0)), gpu_lib.lattice(lat.da_ring(0, extra_quads=True)[:322])]
    B = abi.Batch(gpu_lib, mixed)
    both = np.concatenate([beam, beam])
    d2 = DeviceArray(both)
    with pytest.raises(abi.AccpError, match="differ in structure"):
        B.track(d2.ptr, np.array([0, 64, 128], dtype=np.int64), 1, tuning=spec())
    assert np.array_equal(d2.numpy(), both)


def test_specialized_batched_64_seeds_full_size(gpu_lib, oracle):
    """The batched DA workload at full size (64 seeds, 274,385 points, 1024
    turns, one captured graph) on the specialized kernel.  (1) Bit for bit
    the specialized plain calls per seed.  (2) After 64 turns, on the seeds
    with the smallest and largest grids and two more, the oracle's survival
    maps except within one grid cell of the boundary.  (3) After 1024 turns,
    against the table walker: the same map up to a few hundredths of a
    percent of the points.  At 1024 turns whole patches of this uncorrected
    lattice are chaotic: the oracle loses all nine points of a 3 x 3 patch
    (seed 23, delta 0, around x index 33, y index 6) that the walker keeps,
    and one-ulp changes of the centre point's x or y flip its outcome, so
    rounding decides survival there and no boundary rule holds."""
    from pyaccp.hipmem import DeviceArray, synchronize

    elems, beam, offs = lat.batched_workload(64)
    lats = [gpu_lib.lattice(s) for s in elems]
    B = abi.Batch(gpu_lib, lats)
    t = spec(64, 1, 0)
    d = DeviceArray(beam)
    B.capture(d.ptr, offs, 1024, tuning=t)
    B.replay()
    synchronize()
    sp = d.numpy()
    plain = np.concatenate([L.track(beam[offs[s]:offs[s + 1]], 1024, tuning=t) for s, L in enumerate(lats)])
    assert same_bits(sp, plain)

    sizes = np.diff(offs)
    d64 = DeviceArray(beam)
    B.track(d64.ptr, offs, 64, tuning=t)
    synchronize()
    t64 = d64.numpy()
    for s in sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 0, 37}):
        sl = slice(offs[s], offs[s + 1])
        nx, ny = lat.seed_grid_shape(s)
        so = ~np.isnan(oracle.track(elems[s], beam[sl], 64)[:, 0]).reshape(5, ny, nx)
        sg = ~np.isnan(t64[sl][:, 0]).reshape(5, ny, nx)
        for dd, iy, ix in np.argwhere(sg != so):
            nb = so[dd, max(iy - 1, 0):iy + 2, max(ix - 1, 0):ix + 2]
            assert nb.any() and not nb.all(), f"seed {s}: 64 turns, survival differs away from the boundary"
        print(f"seed {s}: 64 turns, {int(sg.sum())} survivors, {int((sg != so).sum())} outcomes differ from the oracle")

    dw = DeviceArray(beam)
    B.track(dw.ptr, offs, 1024)
    synchronize()
    walker = dw.numpy()
    ss, sw = ~np.isnan(sp[:, 0]), ~np.isnan(walker[:, 0])
    ndiff = int((ss != sw).sum())
    print("1024 turns, survivors: specialized", int(ss.sum()), "walker", int(sw.sum()), "outcomes differ", ndiff)
    assert ndiff <= 5e-4 * len(beam)
    assert abs(int(ss.sum()) - int(sw.sum())) <= 0.01 * sw.sum()

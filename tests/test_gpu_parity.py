"""HIP tracking path against the oracle on the same inputs (gfx950 only).

Tolerance (the tracking contract): per coordinate |gpu - oracle| <=
1e-12 * max(|gpu|, |oracle|, sigma_c) after one turn, sigma_c the beam RMS
of that coordinate; identical loss (NaN) patterns.  FMA contraction and
OCML vs glibc transcendentals make the two differ at the ulp level."""
import numpy as np
import pytest

import pyaccp.elements as el
import pyaccp.lattices as lat
from helpers import assert_coords_close

pytestmark = pytest.mark.gpu

make_golden = pytest.importorskip("golden.make_golden")


def test_each_element_one_pass(gpu_lib, oracle):
    P = make_golden.probe_particles()
    for name, e in make_golden.element_cases():
        L = gpu_lib.lattice([e])
        out = L.track(P, 1)
        ref = oracle.track([e], P, 1)
        assert_coords_close(out, ref, 1e-12, what=name)


def test_single_term_multipoles(gpu_lib, oracle):
    """StrMPoles whose field is one normal term (quadrupole, sextupole,
    octupole; thick and thin) take the kernel's single-term Horner sum; with
    a skew part or a lower-order term they take the general one.  Each
    against the oracle's general Horner, one pass at 1e-12 and, inside a
    ring, 100 turns at 1e-10."""
    P = make_golden.probe_particles()
    cases = [el.strmpole(0.3, [0.0, 1.5]), el.strmpole(0.2, [0.0, 0.0, -25.0]),
             el.strmpole(0.25, [0.0, 0.0, 0.0, 300.0]), el.strmpole(0.0, [0.0, 0.08]),
             el.strmpole(0.0, [0.0, 0.0, 4.0]), el.strmpole(0.0, [0.0, 0.0, 0.0, -60.0]),
             el.strmpole(0.2, [0.0, 0.0, 20.0], [0.0, 0.0, 1.0]), el.strmpole(0.2, [0.0, 0.3, 20.0]),
             el.strmpole(0.2, [0.0, 0.0, 0.0, 0.0, 500.0])]
    for k, e in enumerate(cases):
        assert_coords_close(gpu_lib.lattice([e]).track(P, 1), oracle.track([e], P, 1), 1e-12, what=f"case {k}")
    # the few particles driven unstable overflow, and an overflowed orbit's
    # inf/NaN pattern depends on the operation order: the same particles stay
    # finite, and those agree
    ring = lat.fodo_ring()
    ring.insert(3, el.strmpole(0.1, [0.0, 0.0, 0.0, 150.0], name="OC"))
    ring.insert(9, el.strmpole(0.0, [0.0, 0.01], name="TQ"))
    beam = lat.gaussian_beam(500, seed=4)
    g, o = gpu_lib.lattice(ring).track(beam, 20), oracle.track(ring, beam, 20)
    fg, fo = np.isfinite(g).all(axis=1), np.isfinite(o).all(axis=1)
    assert np.array_equal(fg, fo) and fg.sum() > 480
    assert_coords_close(g[fg], o[fo], 1e-10, what="ring with single-term multipoles, 20 turns")


def test_ring_one_turn_1000_particles(gpu_lib, oracle):
    ring = lat.fodo_ring()
    beam = lat.gaussian_beam(1000, seed=0)
    L = gpu_lib.lattice(ring)
    err = assert_coords_close(L.track(beam, 1), oracle.track(ring, beam, 1), 1e-12, what="1 turn")
    print("max scaled error after 1 turn:", err)
    assert_coords_close(L.track(beam, 10), oracle.track(ring, beam, 10), 1e-10, what="10 turns")


@pytest.mark.parametrize("block,ppl,staging", [(64, 1, 1), (64, 1, 2), (128, 2, 1), (256, 2, 2), (256, 1, 1)])
def test_launch_shapes_agree(gpu_lib, oracle, block, ppl, staging):
    import pyaccp.abi as abi

    ring = lat.da_ring(1, strength_errors=True)
    beam = lat.gaussian_beam(777, seed=3)
    L = gpu_lib.lattice(ring)
    out = L.track(beam, 2, tuning=abi.Tuning(block, ppl, staging, 0))
    assert_coords_close(out, oracle.track(ring, beam, 2), 1e-11, what=f"{block}/{ppl}/{staging}")


def test_element_range_and_turn_offsets(gpu_lib, oracle):
    ring = lat.da_ring(0, extra_quads=True)
    beam = lat.gaussian_beam(300, seed=5)
    L = gpu_lib.lattice(ring)
    for st, ed in ((0, 1), (17, 100), (100, len(ring)), (5, 5)):
        assert_coords_close(L.track(beam, 1, st, ed), oracle.track(ring, beam, 1, st, ed), 1e-12,
                            what=f"range {st}:{ed}")


def test_da_survival_map(gpu_lib, oracle):
    """Seed-0 DA lattice, 21x11x5 grid, 256 turns: losses identical except
    possibly for points next to the survival boundary (chaotic orbits)."""
    da = lat.da_ring(0)
    grid = lat.da_grid(21, 11)
    L = gpu_lib.lattice(da)
    g = L.track(grid, 256)
    o = oracle.track(da, grid, 256)
    sg = ~np.isnan(g[:, 0]).reshape(5, 11, 21)
    so = ~np.isnan(o[:, 0]).reshape(5, 11, 21)
    diff = np.argwhere(sg != so)
    for d, iy, ix in diff:
        nb = so[d, max(iy - 1, 0):iy + 2, max(ix - 1, 0):ix + 2]
        assert nb.any() and not nb.all(), f"survival differs away from the boundary at {(d, iy, ix)}"
    live = so.ravel() & sg.ravel()
    assert_coords_close(g[live], o[live], 1e-6, what="survivors after 256 turns")


def test_radiation_damping_one_turn(gpu_lib, oracle):
    ring = lat.fodo_ring()
    lat.set_rad(ring, 1)
    beam = lat.gaussian_beam(500, seed=2)
    L = gpu_lib.lattice(ring)
    assert_coords_close(L.track(beam, 1), oracle.track(ring, beam, 1), 1e-12, what="rad flag 1")


def test_quantum_excitation_same_stream(gpu_lib, oracle):
    """Flag 2 draws its normals from the counter-based stream keyed by
    (seed, particle id, turn, element, slice): the oracle restates the
    same stream, so one turn agrees to the tracking tolerance."""
    ring = lat.da_ring(0)
    lat.set_rad(ring, 2)
    beam = lat.gaussian_beam(400, seed=4)
    L = gpu_lib.lattice(ring)
    out = L.track(beam, 1, seed=1234, turn0=7, pid0=100)
    ref = oracle.track(ring, beam, 1, seed=1234, turn0=7, pid0=100)
    assert_coords_close(out, ref, 1e-12, what="rad flag 2")
    other = L.track(beam, 1, seed=1235, turn0=7, pid0=100)
    assert np.abs(other[:, 5] - out[:, 5]).max() > 0


def test_monitor_statistics(gpu_lib, oracle):
    ring = lat.da_ring(0)
    beam = lat.gaussian_beam(2000, seed=6)
    beam[::97, 0] = np.nan
    L = gpu_lib.lattice(ring)
    out, recs = L.track(beam, 9, mon_interval=4)
    assert len(recs) == 3
    r = beam.copy()
    for k, t in enumerate((0, 4, 8)):
        if t:
            r = oracle.track(ring, r, 4, turn0=t - 4)
        o = oracle.beam_stat(r, t)
        assert recs[k][0] == t and recs[k][1] == o[1]
        np.testing.assert_allclose(recs[k][2:], o[2:], rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(gpu_lib.beam_stat(beam)[1:], oracle.beam_stat(beam)[1:], rtol=1e-12, atol=1e-20)


def test_edge_cases(gpu_lib, oracle):
    ring = lat.da_ring(0)
    L = gpu_lib.lattice(ring)
    empty = np.zeros((0, 6))
    assert L.track(empty, 5).shape == (0, 6)
    beam = lat.gaussian_beam(65, seed=9)  # one full wave plus one lane
    assert np.array_equal(L.track(beam, 0), beam)
    lost = beam.copy()
    lost[:, 0] = np.nan
    assert np.array_equal(np.isnan(L.track(lost, 3)), np.isnan(lost))
    assert np.array_equal(L.track(lost, 3)[:, 1:], lost[:, 1:])
    # on the aperture edge exactly: x >= xmax is lost (accel.cpp:350)
    edge = np.zeros((4, 6))
    edge[:, 0] = [0.025, -0.025, 0.0249999999, 0.0]
    edge[3, 2] = 0.012
    ap = [el.aperture(0.0, (-0.025, 0.025), (-0.012, 0.012))]
    assert_coords_close(gpu_lib.lattice(ap).track(edge, 1), oracle.track(ap, edge, 1), 0.0, floor=np.ones(6))


def test_batched_equals_plain_per_seed(gpu_lib):
    import pyaccp.abi as abi
    from pyaccp.hipmem import DeviceArray, synchronize

    seeds = [lat.da_ring(s, strength_errors=True) for s in range(4)]
    sizes = [130, 64, 1, 257]
    beams = [lat.gaussian_beam(n, seed=20 + i) * 20 for i, n in enumerate(sizes)]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    lats = [gpu_lib.lattice(s) for s in seeds]
    t = abi.Tuning(64, 1, 2, 0)
    plain = np.concatenate([L.track(b, 40, tuning=t) for L, b in zip(lats, beams)])
    B = abi.Batch(gpu_lib, lats)
    d = DeviceArray(np.concatenate(beams))
    B.track(d.ptr, offs, 40, tuning=t)
    synchronize()
    batched = d.numpy()
    assert np.array_equal(np.isnan(batched), np.isnan(plain))
    assert np.array_equal(np.nan_to_num(batched), np.nan_to_num(plain))


def test_graph_replays_match_eager(gpu_lib):
    import pyaccp.abi as abi
    from pyaccp.hipmem import DeviceArray, synchronize

    seeds = [lat.da_ring(s) for s in range(3)]
    lats = [gpu_lib.lattice(s) for s in seeds]
    B = abi.Batch(gpu_lib, lats)
    grid = lat.da_grid(9, 5)
    beam = np.concatenate([grid, grid, grid])
    offs = np.array([0, len(grid), 2 * len(grid), 3 * len(grid)], dtype=np.int64)
    d_eager = DeviceArray(beam)
    B.track(d_eager.ptr, offs, 12)
    synchronize()
    want = d_eager.numpy()
    d = DeviceArray(beam)
    B.capture(d.ptr, offs, 12)
    other = gpu_lib.lattice(lat.fodo_ring())
    for rep in range(3):
        d.upload(beam)
        synchronize()
        B.replay()
        synchronize()
        other.track(lat.gaussian_beam(50, seed=rep), 1)  # eager work between replays
        got = d.numpy()
        assert np.array_equal(np.nan_to_num(got), np.nan_to_num(want)), f"replay {rep}"


def test_autotune_picks_a_valid_shape(gpu_lib):
    import pyaccp.abi as abi
    from pyaccp.hipmem import DeviceArray

    lats = [gpu_lib.lattice(lat.da_ring(s)) for s in range(2)]
    B = abi.Batch(gpu_lib, lats)
    grid = lat.da_grid(11, 5)
    beam = np.concatenate([grid, grid])
    offs = np.array([0, len(grid), 2 * len(grid)], dtype=np.int64)
    d = DeviceArray(beam)
    best, ms = B.autotune(d.ptr, offs, 2)
    assert best.block in (64, 128, 256) and best.ppl in (1, 2) and best.staging in (1, 2, 3) and ms > 0
    assert best.staging != 3 or best.ppl == 1  # specialized kernels are tried at one particle per lane
    assert np.array_equal(d.numpy(), beam)  # probes ran on a scratch copy


def test_compaction_and_loss_turns(gpu_lib, oracle):
    """Dropping lost particles between chunks of turns changes nothing in the
    coordinates; the recorded loss turn is the turn (0-based) in which the
    oracle first shows the particle lost (-1: lost on input)."""
    import pyaccp.abi as abi

    da = lat.da_ring(0)
    grid = lat.da_grid(21, 11)
    grid[5, 0] = np.nan
    L = gpu_lib.lattice(da)
    a, lt = L.track(grid, 64, loss_turn=True)
    b = L.track(grid, 64, tuning=abi.Tuning(64, 1, 0, -1))
    c = L.track(grid, 64, tuning=abi.Tuning(128, 2, 2, 4))
    assert np.array_equal(np.nan_to_num(a, nan=9.0), np.nan_to_num(b, nan=9.0))
    assert np.array_equal(np.isnan(a), np.isnan(c))
    want = np.full(len(grid), 64, dtype=np.int32)
    want[5] = -1
    r = grid.copy()
    for t in range(64):
        r = oracle.track(da, r, 1)
        newly = np.isnan(r[:, 0]) & (want == 64)
        want[newly] = t
    assert np.array_equal(lt, want)


def test_matrix_syncrad_crabcav_rfmult_wiggler(gpu_lib, oracle):
    """The elements of the full kernel beyond the tracking set: against the
    reference fixtures (elements2.npz) at 1e-12, and with quantum excitation
    and crab-cavity noise against the oracle on the same Philox stream."""
    from conftest import golden
    from tests.golden.make_golden import elements2_cases
    g = golden("elements2.npz")
    P = g["particles"]
    for name, e in elements2_cases():
        assert_coords_close(gpu_lib.lattice([e]).track(P, 1), g[name], 1e-12, what=name)
    beam = lat.gaussian_beam(500, seed=31)
    noisy = [el.rfmult(0.1, 5e8, 0.2, [1e-4, 2e-3, 0.5], [0.0, 1e-3, 0.2], sigphi=2e-3),
             el.syncrad(0.2, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0, rad=el.RAD_LUMP),
             el.syncrad(0.0, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0, rad=el.RAD_LUMP_QE),
             el.crabcav(0.3, 1.5, 0.5, 1.5e9, 0.3, 3000.0, sigphi=1e-3, sigdvv=1e-2)]
    for e in noisy:
        out = gpu_lib.lattice([e]).track(beam, 3, seed=17, turn0=5, pid0=100)
        ref = oracle.track([e], beam, 3, seed=17, turn0=5, pid0=100)
        assert_coords_close(out, ref, 1e-12, what=e.name)
        assert not np.array_equal(out, oracle.track([e], beam, 3, seed=18, turn0=5, pid0=100))
    # a lumped ring: linear one-turn matrix, crab cavity, lumped radiation
    m = np.eye(6)
    mu = 2 * np.pi * 0.31
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = np.cos(mu), 10 * np.sin(mu), -np.sin(mu) / 10, np.cos(mu)
    ring = [el.matrix(1.0, m.ravel()), el.crabcav(0.0, 0.5, 0.0, 1.5e9, 0.0, 3000.0),
            el.syncrad(0.0, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0, rad=el.RAD_LUMP)]
    assert_coords_close(gpu_lib.lattice(ring).track(beam, 50, seed=3), oracle.track(ring, beam, 50, seed=3),
                        1e-10, what="lumped ring, 50 turns")


def test_noise_against_reference_crn(gpu_lib, oracle):
    """The noise cases of noise.npz on the GPU (CrabCav phase and amplitude
    jitter, RFMult phase jitter, SyncRad excitation, BendMPole quantum
    excitation, a line of three noisy elements): against the compiled
    reference fed the same normals in its draw order, within 1e-6 of the
    noise (the reference's normals come through rand()'s grid), and against
    the oracle on the same Philox stream at 1e-12."""
    from conftest import golden
    from tests.golden.make_noise import noise_cases
    from tests.test_oracle_golden import noise_bound
    g = golden("noise.npz")
    P = g["particles"]
    kw = dict(seed=int(g["seed"][0]), turn0=int(g["turn0"][0]), pid0=int(g["pid0"][0]))
    nt = int(g["turns"][0])
    for name, noisy, quiet in noise_cases():
        r = g[name]
        out = gpu_lib.lattice(noisy).track(P, nt, **kw)
        tol, fin = noise_bound(r, oracle.track(quiet, P, nt, **kw))
        assert np.array_equal(np.isfinite(out[:, 0]), fin), name
        err = np.abs(out - r)[fin].max(axis=0)
        print(f"{name}: GPU - reference {err.max():.2e} (bound {tol.max():.2e})", flush=True)
        assert (err <= tol).all(), (name, err, tol)
        assert_coords_close(out, oracle.track(noisy, P, nt, **kw), 1e-12, what=name)


def test_wigtable(gpu_lib, oracle):
    """WigTable on the GPU against the reference fixtures (wigtable.npz) at
    1e-12: thin and sliced tables, descending and non-uniform grids up to
    101 x 21, particles inside, on the nodes of and beyond the grids; then 30
    turns of a ring with a sliced table against the oracle."""
    from conftest import golden
    from tests.golden.make_golden import wigtable_cases
    g = golden("wigtable.npz")
    P = g["particles"]
    cases = wigtable_cases()
    for name, e in cases:
        assert_coords_close(gpu_lib.lattice([e]).track(P, 1), g[name], 1e-12, what=name)
    # short sliced tables at 0.3 of the fixtures' strength: they move the
    # beam by about a millimetre over the run and lose no particle (longer or
    # stronger insertions break the ring's periodicity enough that the
    # sextupoles drive some particles to overflow)

    def insertion(e):
        xl, yl, xk, yk = el.wigtable_arrays(e)
        return el.wigtable(0.05, 4, xl, yl, 0.3 * xk, 0.3 * yk)
    tables = [insertion(cases[1][1]), insertion(cases[3][1])]
    ring = lat.fodo_ring()[:120] + [tables[0]] + lat.fodo_ring()[120:] + [tables[1]]
    beam = lat.gaussian_beam(300, seed=41)
    assert_coords_close(gpu_lib.lattice(ring).track(beam, 30), oracle.track(ring, beam, 30), 1e-10,
                        what="ring with two kick tables, 30 turns")
    # the library keeps its own copy of each table: the records it returns
    # point at that copy and rebuild the same lattice
    L = gpu_lib.lattice(ring)
    arr = L.get()
    wt = [arr[i] for i in range(len(ring)) if arr[i].type == el.WIGTABLE]
    for got, want in zip(wt, tables):
        for a, b in zip(el.wigtable_arrays(got), el.wigtable_arrays(want)):
            assert np.array_equal(a, b)
    assert_coords_close(gpu_lib.lattice([arr[i] for i in range(len(ring))]).track(beam, 5),
                        oracle.track(ring, beam, 5), 1e-11, what="rebuilt from the returned records")


def test_da_scan_at_full_size(gpu_lib, oracle):
    """The DA scan at its full size (101 x 51 x 5 = 25755 points on the
    seed-0 lattice with errors): the survival map after 256 turns against the
    oracle (identical except next to the survival boundary), then all 2048
    turns on the GPU, where two size-independent properties hold: the
    survivors of 2048 turns are a subset of those of 256 turns, and the
    batched path (one seed, captured graph) gives the plain path's map and
    coordinates bit for bit."""
    import pyaccp.abi as abi
    from pyaccp.hipmem import DeviceArray, synchronize

    da = lat.da_ring(0)
    grid = lat.da_grid(101, 51)
    L = gpu_lib.lattice(da)
    g256 = L.track(grid, 256)
    o256 = oracle.track_omp(da, grid, 256, 16)
    sg = ~np.isnan(g256[:, 0]).reshape(5, 51, 101)
    so = ~np.isnan(o256[:, 0]).reshape(5, 51, 101)
    print("survivors after 256 turns: gpu", sg.sum(), "oracle", so.sum(), "differ", int((sg != so).sum()))
    for d, iy, ix in np.argwhere(sg != so):
        nb = so[d, max(iy - 1, 0):iy + 2, max(ix - 1, 0):ix + 2]
        assert nb.any() and not nb.all(), f"survival differs away from the boundary at {(d, iy, ix)}"
    g2048 = L.track(grid, 2048)
    s2048 = ~np.isnan(g2048[:, 0]).reshape(5, 51, 101)
    assert not (s2048 & ~sg).any() and s2048.sum() > 0
    B = abi.Batch(gpu_lib, [L])
    d = DeviceArray(grid)
    B.capture(d.ptr, np.array([0, len(grid)], dtype=np.int64), 2048)
    B.replay()
    synchronize()
    out = d.numpy()
    assert np.array_equal(np.isnan(out), np.isnan(g2048))
    assert np.array_equal(np.nan_to_num(out), np.nan_to_num(g2048))


def test_fodo_1000_particles_1000_turns(gpu_lib, oracle):
    """The plain tracking workload at its full size: 1000 Gaussian particles,
    1000 turns of the 20-cell FODO ring, against the oracle.  Per turn the
    two agree to round-off (fused multiply-adds on the GPU); over 1000 turns
    the sextupoles amplify that (6e-11 measured), so the bound here is 1e-9
    of each coordinate's scale.  The ring has no aperture: the few particles the
    sextupoles drive unstable overflow (inf or NaN, in either order), so the
    test asks for the same set of particles to stay finite."""
    ring = lat.fodo_ring()
    beam = lat.gaussian_beam(1000, seed=1)
    g = gpu_lib.lattice(ring).track(beam, 1000)
    o = oracle.track_omp(ring, beam, 1000, 16)
    fg, fo = np.isfinite(g).all(axis=1), np.isfinite(o).all(axis=1)
    print("finite after 1000 turns: gpu", fg.sum(), "oracle", fo.sum())
    assert np.array_equal(fg, fo) and fg.sum() > 900
    err = assert_coords_close(g[fg], o[fo], 1e-9, what="1000 turns")
    print("max relative difference after 1000 turns", err)


def da_maps_agree(surv_gpu, surv_ref, shape, what):
    """DA survival maps identical except next to the survival boundary: a
    point whose outcome differs must have both survivors and lost points in
    the reference map within one grid cell of it."""
    sg, so = surv_gpu.reshape(shape), surv_ref.reshape(shape)
    for d, iy, ix in np.argwhere(sg != so):
        nb = so[d, max(iy - 1, 0):iy + 2, max(ix - 1, 0):ix + 2]
        assert nb.any() and not nb.all(), f"{what}: survival differs away from the boundary at {(d, iy, ix)}"
    return int((sg != so).sum())


def test_batched_da_64_seeds_full_size(gpu_lib, oracle):
    """The batched DA workload at its full size: 64 error seeds (Quad K and
    SF/SD strength errors on top of the multipole errors), uneven grids of
    500-8000 points (274,385 in all), 1024 turns in one captured graph.
    (1) Survival maps and coordinates equal 64 separate plain calls bit for
    bit.  (2) On the seeds with the smallest and largest grids and two more,
    the batched launch after one turn has the oracle's survivors, its
    particles that stay finite match the oracle's at 1e-12 per coordinate
    (aline.cpp:359-363) and its aperture loss codes are the oracle's; after
    64 turns its survival maps match the oracle's except within one grid cell
    of the boundary (aline.cpp:607-715).  Most grid points at large
    amplitude overflow within the first turn (the sextupoles of this
    uncorrected ring throw them to infinity before the aperture at the end
    of the ring); which of their coordinates end as inf or NaN depends on
    the order of the operations, so only their loss is compared.  A few
    finite orbits just inside that edge amplify round-off so much that the
    oracle itself moves by up to 5e-12 of a coordinate when its initial x
    and y change by about one ulp (of the coordinate, or of 1 mm on the
    axes); such a particle is held to 1e-12 plus 1000
    times that measured sensitivity (every particle that stays inside the
    dynamic aperture for 64 turns has a sensitivity below 1e-13 and is held
    to 1e-12)."""
    import pyaccp.abi as abi
    from pyaccp.hipmem import DeviceArray, synchronize

    elems, beam, offs = lat.batched_workload(64)
    assert len(beam) == 274_385
    lats = [gpu_lib.lattice(s) for s in elems]
    B = abi.Batch(gpu_lib, lats)
    d = DeviceArray(beam)
    B.capture(d.ptr, offs, 1024)
    B.replay()
    synchronize()
    batched = d.numpy()
    plain = np.concatenate([L.track(beam[offs[s]:offs[s + 1]], 1024) for s, L in enumerate(lats)])
    print("survivors after 1024 turns:", int((~np.isnan(batched[:, 0])).sum()))
    assert np.array_equal(np.isnan(batched), np.isnan(plain))
    assert np.array_equal(np.nan_to_num(batched), np.nan_to_num(plain))

    sizes = np.diff(offs)
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 0, 37})
    d1 = DeviceArray(beam)
    B.track(d1.ptr, offs, 1)
    d64 = DeviceArray(beam)
    B.track(d64.ptr, offs, 64)
    synchronize()
    one, t64 = d1.numpy(), d64.numpy()
    for s in pick:
        sl = slice(offs[s], offs[s + 1])
        nx, ny = lat.seed_grid_shape(s)
        assert sizes[s] == 5 * nx * ny
        o1 = oracle.track(elems[s], beam[sl], 1)
        g1 = one[sl]
        assert np.array_equal(np.isnan(g1[:, 0]), np.isnan(o1[:, 0])), f"seed {s}: survivors after one turn"
        fin = np.isfinite(o1).all(axis=1)
        assert np.array_equal(np.isfinite(g1).all(axis=1), fin), f"seed {s}: finite particles after one turn"
        ulp = beam[sl].copy()
        ulp[:, 0] += 2.0 ** -52 * (np.abs(ulp[:, 0]) + 1e-3)
        ulp[:, 2] += 2.0 ** -52 * (np.abs(ulp[:, 2]) + 1e-3)
        op = oracle.track(elems[s], ulp, 1)
        scale = np.maximum(np.maximum(np.abs(g1[fin]), np.abs(o1[fin])), np.sqrt(np.mean(o1[fin] ** 2, axis=0)))
        sens = np.nan_to_num(np.abs(op[fin] - o1[fin]) / scale, nan=1.0, posinf=1.0)
        rel = np.abs(g1[fin] - o1[fin]) / scale
        assert np.all(rel <= 1e-12 + 1e3 * sens), f"seed {s}: one turn, {int((rel > 1e-12 + 1e3 * sens).sum())} coords"
        well = sens.max(axis=1) <= 1e-13
        err = float(rel[well].max()) if well.any() else 0.0
        assert err <= 1e-12, f"seed {s}: one turn, well-conditioned orbits"
        ap = np.isnan(o1[:, 0]) & np.isfinite(o1[:, 4]) & np.isfinite(g1[:, 4])
        assert np.array_equal(g1[ap, 4], o1[ap, 4]), f"seed {s}: aperture codes"
        o64 = oracle.track_omp(elems[s], beam[sl], 64, 16)
        nd = da_maps_agree(~np.isnan(t64[sl, 0]), ~np.isnan(o64[:, 0]), (5, ny, nx), f"seed {s}, 64 turns")
        print(f"seed {s}: {sizes[s]} points ({nx}x{ny}x5), {int(fin.sum())} finite after one turn, "
              f"{int(well.sum())} well-conditioned: max scaled difference {err:.1e} "
              f"(others {float(rel[~well].max()) if (~well).any() else 0.0:.1e}), "
              f"64-turn survivors gpu {int((~np.isnan(t64[sl, 0])).sum())} "
              f"oracle {int((~np.isnan(o64[:, 0])).sum())}, boundary differences {nd}")

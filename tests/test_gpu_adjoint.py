"""Reverse-mode gradient of the turn map on the GPU against central finite
differences: of the GPU's own forward map and of the oracle's.  Tolerance:
1e-6 relative per knob.  The adjoint is exact up to round-off and does not
depend on a step; the central difference carries O(h^2) truncation and a
round-off error growing as 1/h, and on this lattice (chromaticity ~ +100,
off-momentum particles near the half-integer) the loss is strongly
nonlinear in the knobs, so no single step serves every knob.  The check
scans steps 1e-8 .. 1e-5 (times the knob, at least 1e-9) of a fourth-order
central stencil and requires two adjacent steps of the scan to agree, one
within 1e-6 and the other within 5e-6 (a consistency check in the step: one
lucky step among others far off cannot pass); a wrong adjoint would miss
every step."""
import numpy as np
import pytest

import pyaccp.abi as abi
import pyaccp.elements as el
import pyaccp.lattices as lat

pytestmark = pytest.mark.gpu


def fd_steps(ring, f, i, m):
    base = ring[m[0]].k if f == el.FIELD_K else ring[m[0]].polyB[i]
    return [max(r * abs(base), 1e-9) for r in (1e-8, 3e-8, 1e-7, 3e-7, 1e-6, 3e-6, 1e-5)]


def best_fd(grad, loss_at, steps):
    """Agreement with a fourth-order central stencil at two adjacent steps of
    the scan.  Over adjacent pairs of steps, the pair whose worse member
    agrees best; returns (better relative disagreement, worse one, the
    better member's FD value).  The tests ask for 1e-6 and 5e-6: one lucky
    step among others far off cannot pass.  Initial coordinates take steps
    of 1e-5 to 3e-2 of the coordinate (at least of 1e-6): the loss is
    dominated by z^2 ~ 1e-6, and smaller steps of a transverse momentum of
    a few microradians drown in its round-off."""
    fds = []
    for h in steps:
        fd = (-loss_at(2 * h) + 8 * loss_at(h) - 8 * loss_at(-h) + loss_at(-2 * h)) / (12 * h)
        fds.append((abs(grad - fd) / max(abs(fd), 1e-300), fd))
    a, b = min(zip(fds, fds[1:]), key=lambda ab: max(ab[0][0], ab[1][0]))
    lo, hi = min(a, b), max(a, b)
    return lo[0], hi[0], lo[1]


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
    beam = lat.gaussian_beam(256, seed=1) * 0.5
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    nturn = 24
    out = A.forward(beam, nturn, ckpt_every=16)
    assert np.array_equal(out, L.track(beam, nturn))  # forward is the tracking map
    loss, seed = loss_and_seed(out)
    grad, dr0 = A.backward(seed)
    for k, (name, f, i, m) in enumerate(knobs):
        def loss_at(h):
            return loss_and_seed(gpu_lib.lattice(lat.shift_knob(ring, f, i, m, h)).track(beam, nturn))[0]
        rel, worse, fd = best_fd(grad[k], loss_at, fd_steps(ring, f, i, m))
        print(f"{name}: adjoint {grad[k]:+.12e}  best fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
        assert rel < 1e-6 and worse < 5e-6, name
    # dL/dr0 on a few coordinates: without collective elements particles are
    # independent, so only the moved particle's term of the loss changes and
    # the difference is taken on that term alone (the other 255 would only
    # add round-off)
    for (pi, c) in ((0, 0), (17, 1), (100, 2), (200, 5)):
        def loss_at(h):
            b = beam[pi:pi + 1].copy()
            b[0, c] += h
            return loss_and_seed(L.track(b, nturn))[0]
        steps = [r * max(abs(beam[pi, c]), 1e-6) for r in (1e-5, 3e-5, 1e-4, 3e-4, 1e-3, 3e-3, 1e-2, 3e-2)]
        rel, worse, fd = best_fd(dr0[pi, c], loss_at, steps)
        assert rel < 1e-6 and worse < 5e-6, (pi, c, dr0[pi, c], fd)


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
        def loss_at(h):
            return loss_and_seed(oracle.track(lat.shift_knob(ring, f, i, m, h), beam, 12))[0]
        rel, worse, fd = best_fd(grad[k], loss_at, fd_steps(ring, f, i, m))
        assert rel < 1e-6 and worse < 5e-6, (name, grad[k], fd)


def test_adjoint_rejects_unsupported_lattices(gpu_lib):
    ring = lat.fodo_ring()
    lat.set_rad(ring, 2)  # quantum excitation draws noise
    L = gpu_lib.lattice(ring)
    with pytest.raises(abi.AccpError, match="no reverse map"):
        abi.Adjoint(gpu_lib, L, [])
    L2 = gpu_lib.lattice(lat.fodo_ring())
    A = abi.Adjoint(gpu_lib, L2, [(el.FIELD_K, 0, lat.family(lat.fodo_ring(), "QF"))])
    with pytest.raises(abi.AccpError, match="before accp_adjoint_forward"):
        A.backward(np.zeros((4, 6)))


@pytest.mark.parametrize("case", ["8 knobs", "10 knobs", "long line"])
def test_reverse_modes_agree(gpu_lib, case):
    """The sequential sweep and the segment-Jacobian chain are the same
    linear algebra in a different order: gradients and dL/dr0 agree to
    round-off, with a partial last segment (37 turns, checkpoints every 8),
    a particle lost on input and particles lost during the run.  With 8
    knobs the segment Jacobians take all six columns in one lane; with 10
    (the QF and QD families each split into two knobs) one column per lane;
    the long line (the ring twice, 684 elements) holds a table beyond the
    64 KB the reverse kernels stage in LDS, so they read it from global
    memory."""
    ring = adjoint_lattice()
    if case == "long line":
        ring = ring + [e.copy() for e in ring]
    knobs = lat.adjoint_knobs(ring)
    if case == "10 knobs":
        split = []
        for name, f, i, m in knobs:
            if name in ("QF", "QD"):
                split += [(name + "a", f, i, m[0::2]), (name + "b", f, i, m[1::2])]
            else:
                split.append((name, f, i, m))
        knobs = split
        assert len(knobs) == 10
    beam = lat.gaussian_beam(512, seed=3) * 0.5
    beam[5, 0] = np.nan
    beam[7:12, 0] = 0.03  # far outside the dynamic aperture
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    out = A.forward(beam, 37, ckpt_every=8)
    lost = np.isnan(out[:, 0])
    assert lost[5] and lost[7:12].all() and lost.sum() < 20
    seed = np.where(lost[:, None], 0.0, 2.0 * out)
    res = {}
    for mode in (A.REVERSE_SEQUENTIAL, A.REVERSE_SEGMENT_JACOBIAN):
        A.set_reverse(mode)
        res[mode] = A.backward(seed)
    (g1, d1), (g2, d2) = res[A.REVERSE_SEQUENTIAL], res[A.REVERSE_SEGMENT_JACOBIAN]
    assert np.all(np.abs(g1 - g2) <= 1e-11 * np.abs(g1) + 1e-300), (g1, g2)
    assert np.all(d1[lost] == 0) and np.all(d2[lost] == 0)
    scale = np.abs(d1).max(axis=0)
    assert np.all(np.abs(d1 - d2) <= 1e-11 * scale), np.abs(d1 - d2).max(axis=0) / scale
    with pytest.raises(Exception):
        A.set_reverse(7)


def test_adjoint_jacobian_matches_reference_getm66(gpu_lib):
    """The reference's own numerical derivatives: getm66 differences 12 rays
    at +-1e-8 (aline.cpp:900-926; fixture optics.npz from the compiled
    reference).  Six backward passes with dL/dr_out = e_i on one particle
    give row i of the exact one-turn Jacobian; the central difference
    differs from it by O(delta^2) truncation and 1e-16/1e-8 round-off, so
    1e-7 of the largest entry, at delta = 0 and on the dp = 1e-3 orbit."""
    from conftest import golden
    g = golden("optics.npz")
    ring = lat.fodo_ring()
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(el.FIELD_K, 0, lat.family(ring, "QF"))])
    for dp, ref in ((0.0, g["m66"]), (1e-3, g["m66_dpp"])):
        r0 = np.zeros((1, 6))
        r0[0, 5] = dp
        A.forward(r0, 1, ckpt_every=1)
        J = np.zeros((6, 6))
        for i in range(6):
            lam = np.zeros((1, 6))
            lam[0, i] = 1.0
            J[i] = A.backward(lam)[1][0]
        err = np.abs(J - ref).max() / np.abs(ref).max()
        print(f"dp={dp}: adjoint Jacobian vs reference getm66 {err:.2e}")
        assert err <= 1e-7


def test_combined_function_bends(gpu_lib, oracle):
    """Adjoint through BendMPole with gradient and sextupole curvature terms:
    the knob gradients against central differences of the GPU map and of the
    oracle, the two reverse forms against each other, and six backward
    passes against the oracle's getm66 (the reference's 12-ray difference)."""
    ring = lat.combined_function_ring()
    knobs = lat.adjoint_knobs(ring)
    assert [k[0] for k in knobs] == ["QF", "QD", "SF", "SD"]
    # the ring's chromaticity is large: energy offsets of a few 1e-5 keep
    # every particle through the 12 turns
    beam = lat.gaussian_beam(64, seed=5) * 0.5
    beam[:, 5] *= 0.05
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    nturn = 12
    out = A.forward(beam, nturn, ckpt_every=5)
    assert np.array_equal(out, L.track(beam, nturn), equal_nan=True)
    assert not np.isnan(out).any()
    _, seed = loss_and_seed(out)
    res = {}
    for mode in (A.REVERSE_SEQUENTIAL, A.REVERSE_SEGMENT_JACOBIAN):
        A.set_reverse(mode)
        res[mode] = A.backward(seed)
    (grad, dr0), (g2, d2) = res[A.REVERSE_SEQUENTIAL], res[A.REVERSE_SEGMENT_JACOBIAN]
    assert np.all(np.abs(grad - g2) <= 1e-11 * np.abs(grad)), (grad, g2)
    assert np.all(np.abs(dr0 - d2) <= 1e-11 * np.abs(dr0).max(axis=0))
    for k, (name, f, i, m) in enumerate(knobs):
        for track in (lambda r: gpu_lib.lattice(r).track(beam, nturn), lambda r: oracle.track(r, beam, nturn)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(ring, f, i, m, h)))[0],
                              fd_steps(ring, f, i, m))
            print(f"{name}: adjoint {grad[k]:+.12e}  best fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
            assert rel < 1e-6 and worse < 5e-6, name
    for (pi, c) in ((3, 0), (40, 2), (63, 3)):
        def loss_at(h):
            b = beam[pi:pi + 1].copy()
            b[0, c] += h
            return loss_and_seed(L.track(b, nturn))[0]
        steps = [r * max(abs(beam[pi, c]), 1e-6) for r in (1e-5, 3e-5, 1e-4, 3e-4, 1e-3, 3e-3, 1e-2, 3e-2)]
        rel, worse, fd = best_fd(dr0[pi, c], loss_at, steps)
        assert rel < 1e-6 and worse < 5e-6, (pi, c, dr0[pi, c], fd)
    A.set_reverse(A.REVERSE_SEQUENTIAL)
    ref = oracle.getm66(ring)
    r0 = np.zeros((1, 6))
    A.forward(r0, 1, ckpt_every=1)
    J = np.zeros((6, 6))
    for i in range(6):
        lam = np.zeros((1, 6))
        lam[0, i] = 1.0
        J[i] = A.backward(lam)[1][0]
    err = np.abs(J - ref).max() / np.abs(ref).max()
    print(f"adjoint Jacobian vs oracle getm66 {err:.2e}")
    assert err <= 1e-7


def test_strong_quads_take_the_trigonometric_reverse_map(gpu_lib, oracle):
    """Quads beyond the power-series range (2 L^2 |K| = 15) run the
    trigonometric quad6 forward and reverse maps: a transfer line of two such
    quads and a sextupole, tracked 2 passes; knob gradients against central
    differences of the GPU map and of the oracle (1e-6), the two reverse
    forms against each other.  Quads this strong have a phase advance of
    2.7 rad each and no stable periodic cell: the line magnifies amplitudes
    about 360-fold per pass, so the beam starts at 0.2% of the Gaussian
    core's size."""
    line = [el.drift(0.5, "D"), el.quad(0.5, 30.0, "QS"), el.drift(0.5, "D"), el.quad(0.5, -30.0, "QT"),
            el.drift(0.5, "D"), el.strmpole(0.2, [0.0, 0.0, 40.0], name="SX"), el.drift(0.3, "D")]
    knobs = [("QS", el.FIELD_K, 0, lat.family(line, "QS")), ("QT", el.FIELD_K, 0, lat.family(line, "QT")),
             ("SX", el.FIELD_POLYB, 2, lat.family(line, "SX"))]
    beam = lat.gaussian_beam(32, seed=9) * 0.002
    L = gpu_lib.lattice(line)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    out = A.forward(beam, 2, ckpt_every=1)
    assert not np.isnan(out).any()
    _, seed = loss_and_seed(out)
    res = {}
    for mode in (A.REVERSE_SEQUENTIAL, A.REVERSE_SEGMENT_JACOBIAN):
        A.set_reverse(mode)
        res[mode] = A.backward(seed)
    (grad, dr0), (g2, d2) = res[A.REVERSE_SEQUENTIAL], res[A.REVERSE_SEGMENT_JACOBIAN]
    assert np.all(np.abs(grad - g2) <= 1e-11 * np.abs(grad)), (grad, g2)
    assert np.all(np.abs(dr0 - d2) <= 1e-11 * np.abs(dr0).max(axis=0))
    for k, (name, f, i, m) in enumerate(knobs):
        base = line[m[0]].k if f == el.FIELD_K else line[m[0]].polyB[i]
        steps = [r * abs(base) for r in (1e-8, 1e-7, 1e-6, 1e-5)]
        for track in (lambda r: gpu_lib.lattice(r).track(beam, 2), lambda r: oracle.track(r, beam, 2)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(line, f, i, m, h)))[0], steps)
            print(f"{name}: adjoint {grad[k]:+.12e}  best fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
            assert rel < 1e-6 and worse < 5e-6, name


def test_linear_bends(gpu_lib, oracle):
    """The FODO ring built with the linear Bend (bend6) in place of
    BendMPole: knob gradients against central differences of the GPU map and
    of the oracle (1e-6), and the one-turn Jacobian from six backward passes
    against the oracle's getm66."""
    ring = lat.fodo_ring(bend_type="Bend")
    knobs = lat.adjoint_knobs(ring)
    # transverse amplitudes large enough that the sextupole knobs' gradients
    # (1e-5 of the quadrupoles') stand above the differences' round-off;
    # energy offsets small enough that no particle is lost
    beam = lat.gaussian_beam(64, seed=6) * 2.0
    beam[:, 5] *= 0.1
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    nturn = 12
    out = A.forward(beam, nturn, ckpt_every=5)
    assert not np.isnan(out).any()
    _, seed = loss_and_seed(out)
    grad, dr0 = A.backward(seed)
    for k, (name, f, i, m) in enumerate(knobs):
        for track in (lambda r: gpu_lib.lattice(r).track(beam, nturn), lambda r: oracle.track(r, beam, nturn)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(ring, f, i, m, h)))[0],
                              fd_steps(ring, f, i, m))
            print(f"{name}: adjoint {grad[k]:+.12e}  best fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
            assert rel < 1e-6 and worse < 5e-6, name
    ref = oracle.getm66(ring)
    A.forward(np.zeros((1, 6)), 1, ckpt_every=1)
    J = np.array([A.backward(np.eye(6)[i:i + 1])[1][0] for i in range(6)])
    err = np.abs(J - ref).max() / np.abs(ref).max()
    print(f"adjoint Jacobian vs oracle getm66 {err:.2e}")
    assert err <= 1e-7


def test_zero_strength_quads(gpu_lib, oracle):
    """A Quad with K = 0 runs the reference's drift() quirk (z from
    (px + y^2)^2); the adjoint carries it: the FODO ring with two K = 0 quads
    inserted, knob gradients against central differences of the GPU map and
    of the oracle (1e-6)."""
    ring = lat.fodo_ring()
    ring.insert(2, el.quad(0.2, 0.0, "QZ"))
    ring.insert(40, el.quad(0.2, 0.0, "QZ"))
    knobs = lat.adjoint_knobs(ring)
    beam = lat.gaussian_beam(64, seed=12) * 2.0
    beam[:, 5] *= 0.05
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    out = A.forward(beam, 8, ckpt_every=3)
    assert not np.isnan(out).any()
    grad, _ = A.backward(loss_and_seed(out)[1])
    for k, (name, f, i, m) in enumerate(knobs):
        for track in (lambda r: gpu_lib.lattice(r).track(beam, 8), lambda r: oracle.track(r, beam, 8)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(ring, f, i, m, h)))[0],
                              fd_steps(ring, f, i, m))
            print(f"{name}: adjoint {grad[k]:+.12e}  best fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
            assert rel < 1e-6 and worse < 5e-6, name


def test_matrix_element(gpu_lib, oracle):
    """A Matrix element (6x6 map between half drifts, here a weak x-y
    coupling and a small path-length term) inside the FODO ring: the
    adjoint's full variant inverts it by elimination and applies M^T; knob
    gradients against central differences of the GPU map and of the oracle
    (1e-6)."""
    m = np.eye(6)
    m[0, 2], m[2, 0], m[1, 3], m[3, 1], m[4, 0] = 0.01, -0.01, 0.02, -0.02, 1e-3
    ring = lat.fodo_ring()
    ring.insert(5, el.matrix(0.2, m.ravel()))
    knobs = lat.adjoint_knobs(ring)
    beam = lat.gaussian_beam(64, seed=14) * 2.0
    beam[:, 5] *= 0.05
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, mm) for _, f, i, mm in knobs])
    out = A.forward(beam, 8, ckpt_every=3)
    assert not np.isnan(out).any()
    grad, _ = A.backward(loss_and_seed(out)[1])
    for k, (name, f, i, mm) in enumerate(knobs):
        for track in (lambda r: gpu_lib.lattice(r).track(beam, 8), lambda r: oracle.track(r, beam, 8)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(ring, f, i, mm, h)))[0],
                              fd_steps(ring, f, i, mm))
            print(f"{name}: adjoint {grad[k]:+.12e}  best fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
            assert rel < 1e-6 and worse < 5e-6, name


def test_adjoint_lattice_checks(gpu_lib):
    """What the reverse sweep cannot carry is an error, at create and again at
    forward and backward (the lattice can change in between): a K knob on a
    K = 0 Quad (the reference's drift() does not depend on K and quad6 jumps
    in z at K = 0), a singular Matrix, and radiation switched on after
    create (quantum excitation, which draws noise)."""
    ring = lat.fodo_ring()
    ring.insert(2, el.quad(0.2, 0.0, "QZ"))
    L = gpu_lib.lattice(ring)
    with pytest.raises(abi.AccpError, match="K = 0"):
        abi.Adjoint(gpu_lib, L, [(el.FIELD_K, 0, lat.family(ring, "QZ"))])
    # a knob that moves the quad off K = 0 first is fine
    ring[2].k = 0.05
    abi.Adjoint(gpu_lib, gpu_lib.lattice(ring), [(el.FIELD_K, 0, lat.family(ring, "QZ"))])
    m = np.eye(6)
    m[1] = m[0]  # two equal rows
    bad = lat.fodo_ring()
    bad.insert(5, el.matrix(0.0, m.ravel()))
    with pytest.raises(abi.AccpError, match="singular"):
        abi.Adjoint(gpu_lib, gpu_lib.lattice(bad), [])
    good = lat.fodo_ring()
    Lg = gpu_lib.lattice(good)
    A = abi.Adjoint(gpu_lib, Lg, [(el.FIELD_K, 0, lat.family(good, "QF"))])
    beam = lat.gaussian_beam(8, seed=2)
    out = A.forward(beam, 2, ckpt_every=1)
    A.backward(2 * out)
    rad = [e.copy() for e in good]
    lat.set_rad(rad, 2)
    Lg.update(rad)
    with pytest.raises(abi.AccpError, match="no reverse map"):
        A.backward(2 * out)
    with pytest.raises(abi.AccpError, match="no reverse map"):
        A.forward(beam, 2, ckpt_every=1)
    Lg.update(good)
    out2 = A.forward(beam, 2, ckpt_every=1)
    assert np.array_equal(out2, out)


def test_matrix_with_row_swaps(gpu_lib, oracle):
    """A Matrix whose elimination needs row swaps in the reverse sweep: the
    transverse blocks are section maps with phase advances near 90 degrees
    and beta of 8 and 4 m (the second row leads the first column), with an
    x-y coupling and path-length terms; in a transfer line of quads and a
    sextupole tracked 3 passes.  Knob gradients against central differences
    of the GPU map and of the oracle (1e-6, at two adjacent steps)."""
    def block(mu, beta):
        return np.array([[np.cos(mu), beta * np.sin(mu)], [-np.sin(mu) / beta, np.cos(mu)]])
    m = np.eye(6)
    m[0:2, 0:2] = block(1.52, 8.0)
    m[2:4, 2:4] = block(1.45, 4.0)
    assert abs(m[1, 0]) > abs(m[0, 0]) and abs(m[3, 2]) > abs(m[2, 2])
    m[0, 2], m[2, 0] = 0.02, -0.01
    m[4, 0], m[4, 1], m[4, 5] = 1e-3, 2e-3, -0.01
    line = [el.drift(0.4, "D"), el.quad(0.3, 1.1, "QF"), el.drift(0.3, "D"), el.matrix(0.2, m.ravel()),
            el.drift(0.3, "D"), el.quad(0.3, -1.0, "QD"), el.drift(0.2, "D"),
            el.strmpole(0.2, [0.0, 0.0, 15.0], name="SX"), el.drift(0.2, "D")]
    knobs = [("QF", el.FIELD_K, 0, lat.family(line, "QF")), ("QD", el.FIELD_K, 0, lat.family(line, "QD")),
             ("SX", el.FIELD_POLYB, 2, lat.family(line, "SX"))]
    beam = lat.gaussian_beam(32, seed=19) * 5.0
    L = gpu_lib.lattice(line)
    A = abi.Adjoint(gpu_lib, L, [(f, i, mm) for _, f, i, mm in knobs])
    out = A.forward(beam, 3, ckpt_every=2)
    assert not np.isnan(out).any()
    grad, _ = A.backward(loss_and_seed(out)[1])
    for k, (name, f, i, mm) in enumerate(knobs):
        for track in (lambda r: gpu_lib.lattice(r).track(beam, 3), lambda r: oracle.track(r, beam, 3)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(line, f, i, mm, h)))[0],
                              fd_steps(line, f, i, mm))
            print(f"{name}: adjoint {grad[k]:+.12e}  fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
            assert rel < 1e-6 and worse < 5e-6, name


def test_multipole_index_up_to_max_npole(gpu_lib, oracle):
    """The reference's StrMPole kick always sums to order MAX_MPOLE-1
    (accel.cpp:826-839) and its setfield accepts PolyB[MaxNPole]
    (accel.cpp:773-783); its bend kick reads PolyB[1] whatever MaxNPole is
    (atpass.cpp:211-217).  The GPU tracks such coefficients as the oracle
    does (1e-12 per coordinate), and a knob on PolyB[MaxNPole] of a
    sextupole family gets the gradient of that map (1e-6 against central
    differences of the GPU map and the oracle)."""
    beam = lat.gaussian_beam(64, seed=21) * 20.0
    beam[:, 5] *= 0.02
    for e in (el.strmpole(0.2, [0.0, 0.0, 20.0, 300.0], max_npole=3),
              el.strmpole(0.0, [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 4e3], max_npole=2),
              el.bendmpole(1.5, 0.15, polyB=[0.0], name="B1")):
        if e.type == el.BENDMPOLE:
            e.polyB[1] = -0.2
        assert_close = __import__("helpers").assert_coords_close
        assert_close(gpu_lib.lattice([e]).track(beam, 1), oracle.track([e], beam, 1), 1e-12, what=e.name)
    ring = lat.fodo_ring()
    knobs = [("SF3", el.FIELD_POLYB, 3, lat.family(ring, "SF")), ("QF", el.FIELD_K, 0, lat.family(ring, "QF"))]
    assert ring[knobs[0][3][0]].max_npole == 3
    b = lat.gaussian_beam(64, seed=22) * 3.0
    b[:, 5] *= 0.05
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, mm) for _, f, i, mm in knobs])
    out = A.forward(b, 6, ckpt_every=4)
    assert not np.isnan(out).any()
    grad, _ = A.backward(loss_and_seed(out)[1])
    assert grad[0] != 0
    for k, (name, f, i, mm) in enumerate(knobs):
        steps = [0.1, 0.3, 1.0, 3.0] if name == "SF3" else fd_steps(ring, f, i, mm)
        for track in (lambda r: gpu_lib.lattice(r).track(b, 6), lambda r: oracle.track(r, b, 6)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(ring, f, i, mm, h)))[0], steps)
            print(f"{name}: adjoint {grad[k]:+.12e}  fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
            assert rel < 1e-6 and worse < 5e-6, name


def radiating_rf_ring():
    """The FODO ring with classical radiation in its bends and, between its
    cells, a thin crab cavity, a thick RF multipole, lumped damping, a
    horizontal ideal wiggler with its classical loss, a vertical one without,
    a sliced kick table (a third of the wigtable.npz table), and an RF
    cavity: every deterministic single-particle map whose delta changes
    inside the element, or whose kick comes from a table."""
    from tests.golden.make_golden import wigtable_cases
    ring = lat.fodo_ring()
    lat.set_rad(ring, 1)
    xl, yl, xk, yk = el.wigtable_arrays(wigtable_cases()[1][1])
    ins = [(1, el.crabcav(0.0, 0.5, 0.2, 1.5e9, 0.0, 3000.0, name="CRAB")),
           (30, el.rfmult(0.1, 5e8, 0.2, [1e-5, 2e-3, 0.5], [0.0, 1e-3, 0.2], energy=3000.0, name="RFM")),
           (60, el.syncrad(0.2, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0, rad=el.RAD_LUMP_DAMP)),
           (90, el.idealwig(1.0, 10, 0.1, 0.1, 0, rad=el.RAD_DAMP, name="WH")),
           (120, el.idealwig(0.6, 6, 0.08, 0.08, 1, name="WV")),
           (150, el.wigtable(0.05, 4, xl, yl, 0.3 * xk, 0.3 * yk))]
    for i, e in reversed(ins):
        ring.insert(i, e)
    ring.append(el.cavity(0.0, 1.5, lat.RF_FREQ, 0.0, el.DEFAULT_E0, on=1, name="RF"))
    return ring


def test_radiation_rf_wigglers_and_tables(gpu_lib, oracle):
    """The reverse maps of the elements that move delta inside them
    (radiating BendMPole, SyncRad damping, CrabCav and RFMult without noise,
    IdealWig with its loss) and of WigTable, in one ring: forward equals
    tracking, the two reverse forms agree, knob gradients and dL/dr0 against
    central differences of the GPU map and of the oracle (1e-6 at two
    adjacent steps), and the one-turn Jacobian from six backward passes
    against the oracle's getm66 (the reference's 12-ray difference).  Noise,
    quantum excitation and wakes stay errors."""
    ring = radiating_rf_ring()
    knobs = lat.adjoint_knobs(ring)
    assert [k[0] for k in knobs] == ["QF", "QD", "SF", "SD"]
    beam = lat.gaussian_beam(64, seed=7) * 2.0
    beam[:, 5] *= 0.05
    L = gpu_lib.lattice(ring)
    A = abi.Adjoint(gpu_lib, L, [(f, i, m) for _, f, i, m in knobs])
    nturn = 8
    out = A.forward(beam, nturn, ckpt_every=3)
    assert np.array_equal(out, L.track(beam, nturn), equal_nan=True)
    assert not np.isnan(out).any()
    _, seed = loss_and_seed(out)
    res = {}
    for mode in (A.REVERSE_SEQUENTIAL, A.REVERSE_SEGMENT_JACOBIAN):
        A.set_reverse(mode)
        res[mode] = A.backward(seed)
    (grad, dr0), (g2, d2) = res[A.REVERSE_SEQUENTIAL], res[A.REVERSE_SEGMENT_JACOBIAN]
    assert np.all(np.abs(grad - g2) <= 1e-11 * np.abs(grad)), (grad, g2)
    assert np.all(np.abs(dr0 - d2) <= 1e-11 * np.abs(dr0).max(axis=0))
    for k, (name, f, i, m) in enumerate(knobs):
        for track in (lambda r: gpu_lib.lattice(r).track(beam, nturn), lambda r: oracle.track(r, beam, nturn)):
            rel, worse, fd = best_fd(grad[k], lambda h: loss_and_seed(track(lat.shift_knob(ring, f, i, m, h)))[0],
                                     fd_steps(ring, f, i, m))
            print(f"{name}: adjoint {grad[k]:+.12e}  fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}", flush=True)
            assert rel < 1e-6 and worse < 5e-6, name
    for (pi, c) in ((2, 0), (30, 1), (41, 2), (63, 4), (50, 5)):
        def loss_at(h):
            b = beam[pi:pi + 1].copy()
            b[0, c] += h
            return loss_and_seed(oracle.track(ring, b, nturn))[0]
        steps = [r * max(abs(beam[pi, c]), 1e-6) for r in (1e-5, 3e-5, 1e-4, 3e-4, 1e-3, 3e-3, 1e-2, 3e-2)]
        rel, worse, fd = best_fd(dr0[pi, c], loss_at, steps)
        print(f"dL/dr0[{pi},{c}]: adjoint {dr0[pi, c]:+.12e}  fd {fd:+.12e}  rel {rel:.1e} / {worse:.1e}")
        assert rel < 1e-6 and worse < 5e-6, (pi, c, dr0[pi, c], fd)
    A.set_reverse(A.REVERSE_SEQUENTIAL)
    ref = oracle.getm66(ring)
    A.forward(np.zeros((1, 6)), 1, ckpt_every=1)
    J = np.array([A.backward(np.eye(6)[i:i + 1])[1][0] for i in range(6)])
    err = np.abs(J - ref).max() / np.abs(ref).max()
    print(f"adjoint Jacobian vs oracle getm66 {err:.2e}")
    assert err <= 1e-7
    # what has no reverse map
    for e in (el.crabcav(0.0, 0.5, 0.2, 1.5e9, 0.0, 3000.0, sigphi=1e-3),
              el.rfmult(0.0, 5e8, 0.2, [1e-5], [0.0], sigphi=1e-3),
              el.syncrad(0.0, 3000.0, 0.75, 1.3, 1.0, 5e-9, 5e-11, 8e-4, 10.0, 5.0, rad=el.RAD_LUMP_QE)):
        with pytest.raises(abi.AccpError, match="no reverse map"):
            abi.Adjoint(gpu_lib, gpu_lib.lattice(lat.fodo_ring() + [e]), [])
    qe = lat.fodo_ring()
    lat.set_rad(qe, 2)
    with pytest.raises(abi.AccpError, match="no reverse map"):
        abi.Adjoint(gpu_lib, gpu_lib.lattice(qe), [])


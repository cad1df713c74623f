"""Benchmark: batched multi-seed dynamic-aperture tracking on one MI355X.

Workload (the batched DA configuration): 64 error seeds of the DA lattice
(20-cell FODO ring + RF cavity + aperture + seeded b3..b10/a3..a10 thin
multipoles, plus per-seed Quad K and SF/SD PolyB[2] errors), each with an
uneven (x, y) grid of 500..8000 particles at 5 momentum offsets, 1024 turns
in FP64.  One step = one full tracking run of all seeds (one captured HIP
graph: chunked turns with working-set compaction), inputs resident in HBM.

Metric: particle-turns/s = launched particles x turns / wall time of the
step (lost particles count for every requested turn, the way the tracking
call is specified); `alive_particle_turns_per_s` counts only turns a
particle was alive.  Printed beside it: the plain per-seed GPU path (64
sequential single-lattice calls) with its survival maps checked against
the batched run, the roofline of the tracking kernel (algorithmic FP64
flops of the turns actually tracked / kernel time), and the reference CPU
tracking timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "accel-plus_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)

# HBM bytes per k_track launch of the default batched workload (64 seeds,
# 1024 turns, 12 launches per step), per launch shape: rocprofv3 FETCH_SIZE
# and WRITE_SIZE passes, scaled by what each counter reads for a plain device
# copy of known size; profiles/k_track_batched_da_64seeds_1024turns.md.  These
# are constants of that profile, not measured by this run (the counters need
# rocprofv3 around the process); other shapes and workloads report null.
TRAFFIC_BYTES_PER_LAUNCH = {"block=64,ppl=1,lds": 8.99e6, "block=64,ppl=1,specialized": 9.20e6}
TRAFFIC_SOURCE = "profile constant: profiles/k_track_batched_da_64seeds_1024turns.md (not measured in this run)"

# Executed FP64 work of the saturated leg's kernels from rocprofv3 counters
# (SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 per wave x 64 lanes, FMA = 2 flops),
# per particle-turn of the seed-0 DA lattice, per launch shape;
# profiles/k_track_saturated_262144x16.md.
EXECUTED_FP64_FLOPS_PER_PARTICLE_TURN = {"block=256,ppl=1,lds": 66869.0, "block=256,ppl=1,specialized": 67624.0}
# the same counters over one replay of the default batched workload, per launch shape
EXECUTED_FP64_FLOPS_PER_STEP = {"block=64,ppl=1,specialized": 4.451e11}


def shape_name(t):
    """Key of a launch shape in the profile tables."""
    import pyaccp.abi as abi
    return f"block={t.block},ppl={t.ppl},{abi.STAGE_NAMES.get(t.staging, t.staging)}"


def chunk_launches(nturn, chunk_max=256):
    """Tracking launches of one call: chunks of 1, 2, 4, ... turns up to
    chunk_max, compaction in between."""
    k, c, t = 0, 1, 0
    while t < nturn:
        t += min(c, nturn - t)
        k += 1
        c = min(2 * c, chunk_max)
    return k


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """Host CPU model, logical CPUs, the CPUs this process may use, and the
    thread count the CPU baselines use: OMP_NUM_THREADS when the environment
    sets it (the GPU host gives one GPU a share of 16), else the usable CPUs."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        threads = int(os.environ.get("OMP_NUM_THREADS", "")) or usable
    except ValueError:
        threads = usable
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable,
            "threads": max(1, min(threads, usable))}


def cpu_baseline(lats, beam, offs, nturn, seeds=(0, 1), runs=5):
    """The compiled reference (oracle/_ref; the CPU restatement where it is
    not built) on a bounded sample of the same workload: the grids of the
    first seeds, all nturn turns, OpenMP over the host threads of host_cpu();
    the median of `runs` runs."""
    import numpy as np

    from pyoracle import Reference

    cpu = host_cpu()
    cores = cpu["threads"]
    try:
        ref = Reference()
        kind = "reference"
    except FileNotFoundError:
        ref = None
        kind = "port"
    launched = sum(int(offs[s + 1] - offs[s]) for s in seeds) * nturn
    times, alive = [], 0
    for _ in range(runs):
        elapsed, alive = 0.0, 0
        for s in seeds:
            b = beam[offs[s]:offs[s + 1]]
            if kind == "reference":
                line = ref.load(lats[s])
                line.set_threads(cores)
                t0 = time.perf_counter()
                out = line.track(b, nturn)
                elapsed += time.perf_counter() - t0
            else:
                from pyoracle import Oracle

                t0 = time.perf_counter()
                out = Oracle().track_omp(lats[s], b, nturn, cores)
                elapsed += time.perf_counter() - t0
            alive += int((~np.isnan(out[:, 0])).sum())
        times.append(elapsed)
    med = float(np.median(times))
    return {
        "value": launched / med,
        "unit": "particle-turns/s",
        "cores": cores,
        "kind": kind,
        "sample": f"seeds {', '.join(map(str, seeds))} of the batched DA workload ({launched // nturn} particles), "
                  f"{nturn} turns each, ALine::pass(r,N,nturn) with set_max_num_threads={cores}; median of "
                  f"{runs} runs ({', '.join(f'{t:.2f}' for t in times)} s); {alive} particles survive",
        "seconds": round(med, 3),
        "host": cpu,
    }


def config_legs(lib, stream):
    """The plain workloads on the GPU, one tracking call each, inputs in HBM:
    the 20-cell FODO ring with 1000 Gaussian particles for 1000 turns, and
    the full DA scan of the seed-0 lattice (101 x 51 x 5 = 25,755 points,
    2048 turns), each on the kernel specialized for its lattice (64-lane
    workgroups: both end as a few waves' instruction streams) with the
    table-walking kernel beside it; the reference CPU tracking of the same
    FODO beam, and of every tenth DA grid point (median of 3)."""
    import numpy as np

    import pyaccp.abi as abi
    import pyaccp.lattices as lat
    from pyaccp.hipmem import DeviceArray, Event

    log("plain FODO and DA-scan legs")
    res = {}
    cpu = host_cpu()
    try:
        from pyoracle import Reference
        ref = Reference()
    except FileNotFoundError:
        ref = None
    for name, ring, beam, nturn, cpu_stride in (
            ("fodo_1000x1000", lat.fodo_ring(), lat.gaussian_beam(1000, seed=1), 1000, 1),
            ("da_scan_25755x2048", lat.da_ring(0), lat.da_grid(101, 51), 2048, 10)):
        L = lib.lattice(ring)
        d0 = DeviceArray(beam)
        d = DeviceArray(beam)
        e0, e1 = Event(), Event()
        times = {}
        for kernel, tuning in (("specialized", abi.Tuning(64, 1, abi.STAGE_SPECIALIZED, 0)), ("table walker", None)):
            L.track_device(d.ptr, len(beam), 1, stream=stream.handle, tuning=tuning)  # compiles the specialized one
            ms = []
            for _ in range(3):
                d.copy_from(d0, stream.handle)
                e0.record(stream.handle)
                L.track_device(d.ptr, len(beam), nturn, stream=stream.handle, tuning=tuning)
                e1.record(stream.handle)
                ms.append(e0.elapsed_ms(e1))
            times[kernel] = (float(np.median(ms)), d.numpy())
        gms, out = times["specialized"]
        entry = {"particles": len(beam), "turns": nturn, "elements": len(ring), "ms": round(gms, 3),
                 "particle_turns_per_s": len(beam) * nturn / (gms * 1e-3),
                 "survivors": int((~np.isnan(out[:, 0])).sum()), "kernel": "specialized, 64 lanes",
                 "table_walker": {"ms": round(times["table walker"][0], 3),
                                  "survivors": int((~np.isnan(times["table walker"][1][:, 0])).sum())}}
        if ref is not None:
            sample = beam[::cpu_stride]
            line = ref.load(ring)
            line.set_threads(cpu["threads"])
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                line.track(sample, nturn)
                ts.append(time.perf_counter() - t0)
            cs = float(np.median(ts))
            entry["cpu_reference"] = {"particle_turns_per_s": len(sample) * nturn / cs, "cores": cpu["threads"],
                                      "sample": f"{len(sample)} particles (every {cpu_stride}th), {nturn} turns, "
                                                f"median of 3: {cs:.2f} s"}
        res[name] = entry
    return res


def saturated_leg(lib, glat, elems, tuning, stream):
    """The same kernel at full occupancy: a Gaussian core beam that survives
    (2^18 particles on the seed-0 lattice, 16 turns), in 256-lane workgroups
    that share one LDS copy of the table per four waves (the batched leg's
    autotuned shape suits its sparse tail: a 64-lane pick would hold a 63 KB
    table per wave and cap a full GPU at two waves per CU)."""
    import numpy as np

    import pyaccp.flops as flops
    import pyaccp.lattices as lat
    from pyaccp.hipmem import DeviceArray, Event

    log(f"saturated leg ({shape_name(tuning)})")
    nsat, tsat = 1 << 18, 16
    sat_beam = lat.gaussian_beam(nsat, seed=11)
    d_sat0 = DeviceArray(sat_beam)
    d_sat = DeviceArray(sat_beam)
    e0, e1 = Event(), Event()
    sat_ms = []
    for rep in range(3):
        d_sat.copy_from(d_sat0, stream.handle)
        e0.record(stream.handle)
        glat.track_device(d_sat.ptr, nsat, tsat, tuning=tuning, stream=stream.handle)
        e1.record(stream.handle)
        sat_ms.append(e0.elapsed_ms(e1))
    sat_out = d_sat.numpy()
    sat_alive = int((~np.isnan(sat_out[:, 0])).sum())
    sat_fl = nsat * tsat * flops.flops_per_turn(elems, "useful")
    sat_best = min(sat_ms[1:])
    res = {
        "workload": f"{nsat} Gaussian core particles, seed-0 DA lattice, {tsat} turns ({sat_alive} survive)",
        "ms": round(sat_best, 3),
        "particle_turns_per_s": nsat * tsat / (sat_best * 1e-3),
        "achieved_tflops": round(sat_fl / (sat_best * 1e-3) / 1e12, 3),
        "achieved_counts": "algorithmic flops (the reference's formulas, useful terms) / kernel time; the "
                           "kernel executes about half of them (per-element constants formed once, regrouped "
                           "slices), so this rate is no fraction of the FP64 peak: executed_frac_of_fp64_peak "
                           "is",
        "tuning": tuning.as_dict(),
    }
    per_pt = EXECUTED_FP64_FLOPS_PER_PARTICLE_TURN.get(shape_name(tuning))
    if per_pt:
        ex = nsat * tsat * per_pt / (sat_best * 1e-3) / 1e12
        res["executed_tflops"] = round(ex, 3)
        res["executed_frac_of_fp64_peak"] = round(ex / FP64_PEAK_TFLOPS, 4)
        res["executed_source"] = ("FP64 VALU instructions per particle-turn from rocprofv3 counters "
                                  "(profiles/k_track_saturated_262144x16.md) / this run's kernel time")
    return res


def collective_leg(lib, stream, nturn=200, eq_turns=20000):
    """Collective tracking (radiation with quantum excitation, cavity at the
    synchronous phase from the GPU's setrfphase, longitudinal resonator wake
    of 5 kOhm, Q = 1, 10 GHz, 2 nC in 256 bins over +-15 cm) of 1e5 Gaussian particles on
    the seed-0 DA lattice; every turn splits at the wake for the histogram,
    convolution and kick.  Timed over nturn turns on the table-walking kernel
    and on kernels specialized for each segment of the turn; then the whole
    workload (eq_turns turns) on the faster, for the equilibrium energy
    spread, bunch length and emittance."""
    import numpy as np

    import pyaccp.abi as abi
    import pyaccp.elements as el
    import pyaccp.flops as flops
    import pyaccp.lattices as lat
    from pyaccp.hipmem import DeviceArray, Event

    log("collective leg")
    # the window spans +-6 equilibrium bunch lengths (sigma_z ~ 25 mm on this lattice)
    ring = lat.da_ring(0) + [el.longwake(5e3, 1.0, 10e9, 2e-9, 256, -0.15, 0.15, energy=3000.0, on=1)]
    L = lib.lattice(ring)
    phis = L.setrfphase()
    arr = L.get()
    for i in range(len(arr)):
        if arr[i].type == el.BENDMPOLE:
            arr[i].rad = el.RAD_QE
        if arr[i].type in (el.CAVITY, el.LONGWAKE):
            arr[i].on = 1
    L.update(arr)
    n = 100000
    beam = lat.gaussian_beam(n, seed=13)
    d0 = DeviceArray(beam)
    d = DeviceArray(beam)
    e0, e1 = Event(), Event()
    times = {}
    for kernel, tuning in (("specialized", abi.Tuning(256, 1, abi.STAGE_SPECIALIZED, 0)), ("table walker", None)):
        d.copy_from(d0, stream.handle)
        L.track_device(d.ptr, n, 10, seed=3, stream=stream.handle, tuning=tuning)  # warmup, compilation
        d.copy_from(d0, stream.handle)
        e0.record(stream.handle)
        L.track_device(d.ptr, n, nturn, seed=3, stream=stream.handle, tuning=tuning)
        e1.record(stream.handle)
        times[kernel] = (e0.elapsed_ms(e1), d.numpy(), tuning)
    kernel = min(times, key=lambda k: times[k][0])
    ms, out, tuning = times[kernel]
    alive = int((~np.isnan(out[:, 0])).sum())
    fl = n * nturn * flops.flops_per_turn([arr[i] for i in range(len(arr))], "useful")
    res = {
        "workload": f"{n} particles, RADQE + cavity (phi_s = {phis:.6f} from setrfphase) + LongWake "
                    f"(256 bins), {len(ring)} elements, {nturn} turns ({alive} survive)",
        "ms": round(ms, 3),
        "ms_per_turn": round(ms / nturn, 4),
        "particle_turns_per_s": n * nturn / (ms * 1e-3),
        "achieved_tflops": round(fl / (ms * 1e-3) / 1e12, 3),
        "kernel": kernel,
        "ms_per_turn_by_kernel": {k: round(v[0] / nturn, 4) for k, v in times.items()},
    }
    if eq_turns:
        # the whole collective workload: 20000 turns (five longitudinal
        # damping times, U0 = 0.75 MeV at 3 GeV), in pieces of 250 turns that
        # continue one RNG stream, a record after each piece
        log("collective leg: equilibrium run")
        beta, alpha, eta, etap = start_optics(lib, ring)
        d.copy_from(d0, stream.handle)
        piece = 250
        recs = []
        t0 = time.perf_counter()
        for k in range(eq_turns // piece):
            L.track_device(d.ptr, n, piece, seed=3, turn0=k * piece, stream=stream.handle, tuning=tuning)
            r = d.numpy()
            recs.append([(k + 1) * piece] + beam_record(lib.beam_stat(r), r, beta, alpha, eta, etap))
        sec = time.perf_counter() - t0
        recs = np.array(recs)
        cols = ("survivors", "sigma_delta", "sigma_z_m", "emittance_x_projected_m_rad", "emittance_x_betatron_m_rad",
                "emittance_x_median_action_m_rad")

        def means(t_lo, t_hi):
            sel = (recs[:, 0] > t_lo) & (recs[:, 0] <= t_hi)
            return {c: float(recs[sel, j + 1].mean()) for j, c in enumerate(cols)}
        res["equilibrium"] = {
            "turns": eq_turns,
            "seconds": round(sec, 2),
            "particle_turns_per_s": n * eq_turns / sec,
            "survivors": int(recs[-1, 1]),
            "last_4000_turns": means(eq_turns - 4000, eq_turns),
            "note": "means over the records of the last 4000 turns; the emittances are horizontal, with the "
                    "dispersion (from the one-turn matrix) taken out for the betatron ones; sigma_delta of "
                    "the ring without the wake is 8.3e-4 (Sands, 3 GeV, rho = 9.55 m)",
        }
        fixture = os.path.join(ROOT, "tests", "golden", "collective.npz")
        if os.path.exists(fixture) and eq_turns >= 8000:
            g = np.load(fixture, allow_pickle=False)
            gr = g["records"]
            sel = (gr[:, 0] > 6000) & (gr[:, 0] <= 8000)
            mine = means(6000, 8000)
            ref = {"survivors_fraction": float(gr[sel, 1].mean() / g["n"][0]), "sigma_delta": float(gr[sel, 2].mean()),
                   "sigma_z_m": float(gr[sel, 3].mean()),
                   "emittance_x_median_action_m_rad": float(gr[sel, 8].mean())}
            res["equilibrium"]["turns_6000_8000_vs_oracle"] = {
                "gpu": {"survivors_fraction": mine["survivors"] / n, "sigma_delta": mine["sigma_delta"],
                        "sigma_z_m": mine["sigma_z_m"],
                        "emittance_x_median_action_m_rad": mine["emittance_x_median_action_m_rad"]},
                "oracle": ref,
                "oracle_run": f"tests/golden/collective.npz: the oracle on {int(g['n'][0])} particles of the same "
                              f"Gaussian, the same lattice, wake and phase (another sample of the beam)",
            }
    return res


def tail_leg(lib, stream, n=48, nturn=256):
    """The sparse tail of the batched DA run, isolated: one partly filled
    wave (48 particles inside the dynamic aperture, 64-lane workgroup, table
    in LDS, no compaction) tracking nturn turns of the seed-0 DA lattice, and
    of sub-lattices holding one element class of it (its 40 dipole
    BendMPoles, 40 sextupoles, 80 thin error multipoles, 40 quads, 120
    drifts).  Each costs one wave's instruction stream per turn: the
    per-turn time of the batched run's last launches."""
    import numpy as np

    import pyaccp.abi as abi
    import pyaccp.elements as el
    import pyaccp.lattices as lat
    from pyaccp.hipmem import DeviceArray, Event

    log("tail leg")
    rng = np.random.default_rng(5)
    beam = np.zeros((n, 6))
    beam[:, 0] = rng.uniform(-1e-3, 1e-3, n)
    beam[:, 2] = rng.uniform(0.0, 5e-4, n)
    da = lat.da_ring(0)
    groups = (("full lattice", da),
              ("dipole bends", [e for e in da if e.type == el.BENDMPOLE]),
              ("sextupoles", [e for e in da if e.type == el.STRMPOLE and e.len > 0]),
              ("thin error multipoles", [e for e in da if e.type == el.STRMPOLE and e.len == 0]),
              ("quads", [e for e in da if e.type == el.QUAD]),
              ("drifts", [e for e in da if e.type == el.DRIFT]))
    res = {}
    for kernel, tuning in (("table walker", abi.Tuning(64, 1, abi.STAGE_LDS, -1)),
                           ("specialized", abi.Tuning(64, 1, abi.STAGE_SPECIALIZED, -1))):
        res[kernel] = {}
        for name, elems in groups:
            L = lib.lattice(elems)
            d = DeviceArray(beam)
            # untimed: compiles the specialized kernel of this structure
            L.track_device(d.ptr, n, 4, stream=stream.handle, tuning=tuning)
            d.upload(beam)
            e0, e1 = Event(), Event()
            e0.record(stream.handle)
            L.track_device(d.ptr, n, nturn, stream=stream.handle, tuning=tuning)
            e1.record(stream.handle)
            us = e0.elapsed_ms(e1) / nturn * 1e3
            res[kernel][name] = {"elements": len(elems), "us_per_turn": round(us, 2),
                                 "ns_per_element": round(us / len(elems) * 1e3, 1),
                                 "lost": int(np.isnan(d.numpy()[:, 0]).sum())}
    return {"workload": f"one wave of {n} particles (64-lane workgroup, constants in LDS, no compaction), "
                        f"{nturn} turns; the table-walking kernel and the kernel specialized for each lattice",
            "per_class": res}


def start_optics(lib, ring):
    """beta_x, alpha_x, eta_x, eta_x' at the ring start from the GPU's
    one-turn matrix of the ring with radiation, cavity and wake off."""
    import numpy as np

    import pyaccp.elements as el

    lin = [e.copy() for e in ring]
    for e in lin:
        if e.type == el.BENDMPOLE:
            e.rad = 0
        if e.type in (el.CAVITY, el.LONGWAKE):
            e.on = 0
    m = lib.lattice(lin).getm66()
    eta = np.linalg.solve(np.eye(4) - m[:4, :4], m[:4, 5])
    c = 0.5 * (m[0, 0] + m[1, 1])
    sn = np.copysign(np.sqrt(1.0 - c * c), m[0, 1])
    return m[0, 1] / sn, (m[0, 0] - m[1, 1]) / (2.0 * sn), eta[0], eta[1]


def beam_record(stat, r, beta, alpha, eta, etap):
    """survivors, sigma_delta, sigma_z, projected horizontal emittance, and the
    betatron emittance from the second moments and from the median action
    (eps = median(J) / ln 2 for a Gaussian beam)."""
    import numpy as np

    emit = float(np.sqrt(max(stat[8] ** 2 * stat[9] ** 2 - stat[14] ** 2, 0.0)))
    live = r[~np.isnan(r[:, 0])]
    dd = live[:, 5] - live[:, 5].mean()
    xb = live[:, 0] - live[:, 0].mean() - eta * dd
    pb = live[:, 1] - live[:, 1].mean() - etap * dd
    e_rms = float(np.sqrt(max(np.mean(xb * xb) * np.mean(pb * pb) - np.mean(xb * pb) ** 2, 0.0)))
    gamma = (1.0 + alpha * alpha) / beta
    action = 0.5 * (gamma * xb * xb + 2.0 * alpha * xb * pb + beta * pb * pb)
    return [float(stat[1]), float(stat[13]), float(stat[12]), emit, e_rms, float(np.median(action) / np.log(2.0))]


def adjoint_leg(lib, stream):
    """Adjoint gradient workload: loss = sum of squared final coordinates of
    4096 particles after 512 turns of the DA lattice with 20 extra quads
    (radiation off, cavity on); gradient w.r.t. 8 family knobs, checkpoints
    every 16 turns.  The beam is the Gaussian core at half transverse size
    and a quarter of the bunch length and energy spread: with chromaticity
    near +100, particles beyond about 1e-3 in delta cross strong resonances
    within 512 turns and their finite differences stop converging, so this
    beam keeps the motion regular and finite differences meaningful: the
    cost the adjoint replaces is timed as the 16 forwards of a two-point
    central difference, and the gradient is checked against fourth-order
    central stencils at knob steps 1e-7 .. 1e-4 (relative), agreement
    required of two adjacent steps, all computed with the same GPU forward
    map.  Every tracking pass (the forward, the adjoint's checkpointing and
    recomputation passes, the finite differences) runs on the kernel
    specialized for the lattice (its default launch shape, set with
    accp_lattice_set_tuning); the table walker's forward is reported beside
    it."""
    import numpy as np

    import pyaccp.abi as abi
    import pyaccp.lattices as lat
    from pyaccp.hipmem import DeviceArray, Event, synchronize

    ring = lat.da_ring(0, extra_quads=True)
    lat.set_rad(ring, 0)
    lat.set_cavity(ring, 1)
    knobs = lat.adjoint_knobs(ring)
    beam = lat.gaussian_beam(4096, seed=1) * np.array([0.5, 0.5, 0.5, 0.5, 0.25, 0.25])
    nturn = 512
    L = lib.lattice(ring)
    d0 = DeviceArray(beam)
    d = DeviceArray(beam)
    e0, e1 = Event(), Event()

    def forward_ms(lattice):
        d.copy_from(d0, stream.handle)
        e0.record(stream.handle)
        lattice.track_device(d.ptr, len(beam), nturn, stream=stream.handle)
        e1.record(stream.handle)
        return e0.elapsed_ms(e1)

    forward_ms(L)
    fwd_walker = min(forward_ms(L) for _ in range(3))
    tuning = abi.Tuning(64, 1, abi.STAGE_SPECIALIZED, 0)
    L.set_tuning(tuning)
    forward_ms(L)  # compiles the specialized kernel
    fwd = min(forward_ms(L) for _ in range(3))
    A = abi.Adjoint(lib, L, [(f, i, m) for _, f, i, m in knobs])
    times = []
    for _ in range(3):
        synchronize()
        t0 = time.perf_counter()
        out = A.forward(beam, nturn, ckpt_every=16)
        # loss = sum of squares over the particles that survive (lost ones
        # carry no gradient)
        alive = ~np.isnan(out[:, 0])
        grad, _ = A.backward(np.where(alive[:, None], 2.0 * out, 0.0))
        synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    adj = min(times)
    synchronize()
    t0 = time.perf_counter()

    def loss(lattice):
        d.copy_from(d0, stream.handle)
        lattice.track_device(d.ptr, len(beam), nturn, stream=stream.handle)
        r = d.numpy()[alive]
        return float(np.sum(r * r))

    def shifted(name_f_i_m, h):
        _, f, i, m = name_f_i_m
        sl = lib.lattice(lat.shift_knob(ring, f, i, m, h))
        sl.set_tuning(tuning)
        return sl

    # cost the adjoint replaces: two-point central differences, 16 forwards
    for kn in knobs:
        loss(shifted(kn, 1e-7)) - loss(shifted(kn, -1e-7))
    synchronize()
    fd_ms = (time.perf_counter() - t0) * 1e3

    def knob_value(f, i, m):
        return ring[m[0]].k if f == 1 else ring[m[0]].polyB[i]

    rel, rel_pair = [], []
    for k, kn in enumerate(knobs):
        errs = []
        for rs in (1e-7, 3e-7, 1e-6, 3e-6, 1e-5, 3e-5, 1e-4):
            h = max(rs * abs(knob_value(*kn[1:])), 1e-9)
            fd = (-loss(shifted(kn, 2 * h)) + 8 * loss(shifted(kn, h)) - 8 * loss(shifted(kn, -h))
                  + loss(shifted(kn, -2 * h))) / (12 * h)
            errs.append(abs(grad[k] - fd) / abs(fd))
        # the adjacent pair of steps whose worse member agrees best
        a, b = min(zip(errs, errs[1:]), key=lambda ab: max(ab))
        rel.append(min(a, b))
        rel_pair.append(max(a, b))
    return {
        "workload": f"4096 particles (Gaussian core; x,y at half, z,delta at quarter size) x {nturn} turns, {len(ring)} elements, "
                    f"{len(knobs)} knobs (QF,QD K; SF,SD PolyB[2]; QM1-4 K), checkpoints every 16 turns",
        "forward_ms": round(fwd, 3),
        "forward_ms_by_kernel": {"specialized": round(fwd, 3), "table walker": round(fwd_walker, 3)},
        "adjoint_ms": round(adj, 3),
        "adjoint_over_forward": round(adj / fwd, 3),
        "fd_central_16_forwards_ms": round(fd_ms, 3),
        "fd_over_adjoint": round(fd_ms / adj, 3),
        "max_rel_diff_vs_fd": max(rel),
        "rel_diff_vs_fd": [float(x) for x in rel],
        "rel_diff_vs_fd_adjacent_step": [float(x) for x in rel_pair],
        "fd_check": "fourth-order central stencils at knob steps 1e-7, 3e-7, 1e-6, 3e-6, 1e-5, 3e-5, 1e-4 (relative): "
                    "the adjacent pair of steps whose worse member agrees best; both members reported",
        "survivors": int(alive.sum()),
        "knobs": [k[0] for k in knobs],
        "gradient": [float(g) for g in grad],
    }


def diffusion_leg(lib, nx=256, ny=128, nturn=1024, cpu_sample=32):
    """Frequency map (trackDiffusion, aline.cpp:1421-1484) of the seed-0 DA
    lattice with radiation and cavities off: a 256 x 128 grid of starting
    points (x0 to 8 mm, y0 to 4 mm), 1024 turns, NAFF of both halves of
    each plane; the whole call (one tracking launch sequence recording every
    turn, one NAFF launch).  The reference's own trackDiffusion, one particle
    after another on one core, on the first cpu_sample points of the grid.
    The tracking runs on the kernel specialized for the lattice (its default
    launch shape); the table walker's time is reported beside it."""
    import numpy as np

    import pyaccp.abi as abi
    import pyaccp.elements as el
    import pyaccp.lattices as lat

    log("diffusion leg")
    ring = []
    for e in lat.da_ring(0):
        e = e.copy()
        if e.type == el.CAVITY:
            e.on = 0
        if e.type == el.BENDMPOLE:
            e.rad = 0
        ring.append(e)
    L = lib.lattice(ring)
    X, Y = np.meshgrid(np.linspace(1e-4, 8e-3, nx), np.linspace(1e-4, 4e-3, ny))
    x0, y0 = X.ravel(), Y.ravel()
    walker_s = None
    for tuning in (None, abi.Tuning(64, 1, abi.STAGE_SPECIALIZED, 0)):
        L.set_tuning(tuning)
        L.trackdiffusion(64, x0[:8], y0[:8], 0.05, 0.45, 0.05, 0.45)  # warmup, compilation
        t0 = time.perf_counter()
        tot, nux, nuy, dif = L.trackdiffusion(nturn, x0, y0, 0.05, 0.45, 0.05, 0.45)
        gpu_s = time.perf_counter() - t0
        if tuning is None:
            walker_s = gpu_s
    live = int((~np.isnan(dif)).sum())
    sig = np.cos(2 * np.pi * np.outer(np.linspace(0.1, 0.4, 4 * live), np.arange(nturn // 2)))
    t0 = time.perf_counter()
    lib.naff(sig, 0.05, 0.45)
    naff_s = time.perf_counter() - t0
    res = {
        "workload": f"trackDiffusion of a {nx} x {ny} grid, {nturn} turns, seed-0 DA lattice "
                    f"({len(ring)} elements), {live} of {nx * ny} points survive",
        "ms": round(gpu_s * 1e3, 3),
        "points_per_s": nx * ny / gpu_s,
        "table_walker_ms": round(walker_s * 1e3, 3),
        "naff_ms": round(naff_s * 1e3, 3),
        "naff_signals": 4 * live,
        "naff_samples": nturn // 2,
    }
    try:
        from pyoracle import Reference

        line = Reference().load(ring)
        line.set_threads(1)
        t0 = time.perf_counter()
        _, rnux, rnuy, _ = line.trackdiffusion(nturn, x0[:cpu_sample], y0[:cpu_sample], 0.05, 0.45, 0.05, 0.45)
        cpu_s = time.perf_counter() - t0
        ok = ~np.isnan(rnux)
        res["cpu_reference"] = {
            "points_per_s": cpu_sample / cpu_s,
            "cores": 1,
            "sample": f"the first {cpu_sample} grid points, {cpu_s:.2f} s",
            "max_tune_difference": float(np.max(np.abs(np.concatenate([nux[:cpu_sample][ok] - rnux[ok],
                                                                       nuy[:cpu_sample][ok] - rnuy[ok]]))))
            if ok.any() else None,
        }
        res["speedup_vs_reference_core"] = round((nx * ny / gpu_s) / (cpu_sample / cpu_s), 1)
    except FileNotFoundError:
        pass
    return res


def side_legs(lib, stream, legs, lat0, glat0, args, result):
    """The legs beside the batched one that need only the seed-0 lattice."""
    import pyaccp.abi as abi

    if "saturated" in legs:
        result["saturated"] = saturated_leg(lib, glat0, lat0, abi.Tuning(256, 1, abi.STAGE_LDS, 0), stream)
        result["saturated"]["specialized"] = saturated_leg(
            lib, glat0, lat0, abi.Tuning(256, 1, abi.STAGE_SPECIALIZED, 0), stream)
    if "tail" in legs:
        result["tail"] = tail_leg(lib, stream)
    if "configs" in legs:
        result["configs"] = config_legs(lib, stream)
    if "collective" in legs:
        result["collective"] = collective_leg(lib, stream, eq_turns=args.eq_turns)
    if "adjoint" in legs:
        log("adjoint leg")
        result["adjoint"] = adjoint_leg(lib, stream)
    if "diffusion" in legs:
        result["diffusion"] = diffusion_leg(lib)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nseed", type=int, default=64)
    ap.add_argument("--nturn", type=int, default=1024)
    ap.add_argument("--no-plain", action="store_true", help="skip the per-seed plain comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--tune", default="",
                    help="block,ppl,staging (scalar|lds|specialized) instead of autotuning")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient leg")
    ap.add_argument("--eq-turns", type=int, default=20000,
                    help="turns of the collective leg's equilibrium run (0: skip it)")
    ap.add_argument("--legs", default="saturated,tail,configs,collective,adjoint,diffusion,plain,cpu",
                    help="comma list of the legs run after the batched one "
                         "(saturated, tail, configs, collective, adjoint, diffusion, plain, cpu); 'none' for the "
                         "batched leg alone")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched leg (and its headline fields) and run only the listed legs")
    args = ap.parse_args()
    legs = set() if args.legs == "none" else set(args.legs.split(","))
    unknown = legs - {"saturated", "tail", "configs", "collective", "adjoint", "diffusion", "plain", "cpu"}
    if unknown:
        ap.error(f"unknown legs {sorted(unknown)}")
    for flag, leg in ((args.no_plain, "plain"), (args.no_cpu, "cpu"), (args.no_adjoint, "adjoint")):
        if flag:
            legs.discard(leg)
    if args.eq_turns and args.eq_turns < 4000:
        ap.error("--eq-turns is 0 or at least 4000 (statistics over the last 4000 turns)")
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this backend has no "
              f"multi-GPU path: seeds and particles are independent and would shard with zero exchange.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import pyaccp
    import pyaccp.abi as abi
    import pyaccp.flops as flops
    import pyaccp.lattices as lat
    from pyaccp.hipmem import DeviceArray, Event, Stream, synchronize

    lib = pyaccp.load()
    if lib.device_count() < 1:
        raise SystemExit("no HIP device")
    stream = Stream()

    if args.no_batched:
        if legs & {"plain", "cpu"}:
            ap.error("the plain and cpu legs compare with the batched run (drop --no-batched)")
        result = {"legs": sorted(legs)}
        lat0 = lat.da_ring(0, strength_errors=True)
        side_legs(lib, stream, legs, lat0, lib.lattice(lat0), args, result)
        print(json.dumps(result), flush=True)
        return

    log("building workload")
    lats, beam, offs = lat.batched_workload(args.nseed)
    n = len(beam)
    nturn = args.nturn
    glats = [lib.lattice(x) for x in lats]
    B = abi.Batch(lib, glats)
    d_pristine = DeviceArray(beam)
    d_r = DeviceArray(beam)
    d_lt = DeviceArray(np.zeros(n, dtype=np.int32))

    if args.tune:
        bl, pp, sg = args.tune.split(",")
        tuning = abi.Tuning(int(bl), int(pp), {"scalar": abi.STAGE_SCALAR, "lds": abi.STAGE_LDS,
                                               "specialized": abi.STAGE_SPECIALIZED}[sg], 0)
    else:
        log("autotuning")
        # each shape runs the whole workload once: the sparse tail, not the
        # dense first turns, sets the step time, and a short probe sees only
        # the latter (15 shapes x 2 runs, a few seconds, plus the
        # specialized kernels' compilations)
        tuning, probe_ms = B.autotune(d_r.ptr, offs, nturn)
        log(f"autotune: {tuning} ({probe_ms:.2f} ms for {nturn} turns)")

    B.capture(d_r.ptr, offs, nturn, tuning=tuning, d_loss_turn=d_lt.ptr)

    for _ in range(args.warmup):
        d_r.copy_from(d_pristine, stream.handle)
        B.replay(stream.handle)
    synchronize()
    ev0 = [Event() for _ in range(args.steps)]
    ev1 = [Event() for _ in range(args.steps)]
    log(f"timing {args.steps} steps")
    synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        d_r.copy_from(d_pristine, stream.handle)
        ev0[k].record(stream.handle)
        B.replay(stream.handle)
        ev1[k].record(stream.handle)
    synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([a.elapsed_ms(b) for a, b in zip(ev0, ev1)]))
    ms_per_step = wall / args.steps * 1e3

    out = d_r.numpy()
    lt = d_lt.numpy()
    alive_turns = np.where(lt >= nturn, nturn, np.maximum(lt + 1, 0)).astype(np.int64)
    survived = ~np.isnan(out[:, 0])
    launched_pt = n * nturn
    alive_pt = int(alive_turns.sum())
    # algorithmic flops of the turns actually tracked, per seed lattice
    fl_useful = 0.0
    fl_ref = 0.0
    for s in range(len(lats)):
        at = int(alive_turns[offs[s]:offs[s + 1]].sum())
        fl_useful += at * flops.flops_per_turn(lats[s], "useful")
        fl_ref += at * flops.flops_per_turn(lats[s], "reference")
    launches = chunk_launches(nturn)
    # per launch: algorithmic flops / average launch time = the same ratio
    # over the whole step (the bookkeeping kernels are < 0.1% of it)
    achieved = (fl_useful / launches) / (kernel_ms / launches * 1e-3) / 1e12
    default_workload = args.nseed == 64 and nturn == 1024

    result = {
        "metric": "particle-turns/s FP64 (batched vs plain GPU)",
        "value": launched_pt / (ms_per_step * 1e-3),
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattices and particle grids generated in bench.py)",
        "config": {
            "workload": f"batched DA: {len(lats)} error seeds of the 20-cell FODO DA lattice "
                        f"({len(lats[0])} elements), uneven grids, {n} particles, {nturn} turns, one graph",
            "particles": int(n),
            "turns": nturn,
            "seeds": len(lats),
            "tuning": tuning.as_dict(),
        },
        "alive_particle_turns_per_s": alive_pt / (ms_per_step * 1e-3),
        "alive_particle_turns": alive_pt,
        "survivors": int(survived.sum()),
        "roofline": {
            "bound": "valu",
            "compute": "FP64 vector ALU (no dense contraction, no MFMA; the bound is neither HBM nor MFMA); "
                       "peak = MI355X FP64 78.6 TFLOP/s",
            "achieved": round(achieved, 4),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
            "traffic": TRAFFIC_BYTES_PER_LAUNCH.get(shape_name(tuning)) if default_workload else None,
            "traffic_unit": "HBM bytes per k_track launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "traffic_source": TRAFFIC_SOURCE,
            "launches_per_step": launches,
            "ms_per_launch": round(kernel_ms / launches, 3),
            "flops_per_launch_useful": fl_useful / launches,
            "flops_per_step_useful": fl_useful,
            "flops_per_step_reference_algorithm": fl_ref,
            "kernel_ms": round(kernel_ms, 3),
        },
    }
    shape = shape_name(tuning)
    if default_workload and shape in EXECUTED_FP64_FLOPS_PER_STEP:
        ex = EXECUTED_FP64_FLOPS_PER_STEP[shape] / (kernel_ms * 1e-3) / 1e12
        result["roofline"]["executed_tflops"] = round(ex, 4)
        result["roofline"]["executed_frac"] = round(ex / FP64_PEAK_TFLOPS, 5)
        result["roofline"]["executed_source"] = ("FP64 VALU instructions of one replay from rocprofv3 counters "
                                                 "(profiles/k_track_batched_da_64seeds_1024turns.md) / this "
                                                 "run's kernel time")

    side_legs(lib, stream, legs, lats[0], glats[0], args, result)

    if "plain" in legs:
        # the same kernel and launch shape, called once per seed (no batching,
        # no graph); the first call of a specialized shape compiles nothing
        # here, the batched run already did
        log("plain per-seed GPU path")
        d_p = DeviceArray(beam)
        synchronize()
        t0 = time.perf_counter()
        for s in range(len(lats)):
            cnt = int(offs[s + 1] - offs[s])
            glats[s].track_device(d_p.ptr + int(offs[s]) * 48, cnt, nturn, tuning=tuning, stream=stream.handle)
        synchronize()
        plain_s = time.perf_counter() - t0
        pout = d_p.numpy()
        same = bool(np.array_equal(np.isnan(pout[:, 0]), np.isnan(out[:, 0])))
        result["plain_per_seed"] = {
            "value": launched_pt / plain_s,
            "ms": round(plain_s * 1e3, 3),
            "batched_speedup": round(plain_s * 1e3 / ms_per_step, 3),
            "identical_survival_maps": same,
            "tuning": tuning.as_dict(),
        }

    if "cpu" in legs:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(lats, beam, offs, nturn)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

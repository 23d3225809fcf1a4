"""Benchmark: FP64 particle-turns/s through the synthetic 20,000-element ring.

Workload (one GPU): the 20k-element LHC-like ring of
sixtracklib_amd.lattices.lhc_like_ring(seed=0) and 1,048,576 particles on a
polar grid (16,384 amplitudes in (0, 20] sigma x 64 angles, delta 2.7e-4),
tracked 1000 turns with the HIP tracker (TrackJob, FMA build by default).

The timed region is always the whole 1000-turn job (``--turns``): a step
is one track_until call, and the K steps split the turns evenly (step k
tracks the resident particle set to turn round(1000 (k + 1) / K)), so 10
steps are 100 turns each and 20 steps 50 each.  Warmup steps run first on
the same particles, the first W steps of the same split; the initial state
is then pushed again so the timed steps cover turns 0..1000.  Only tracked
particle-turns count: sum over particles of (final at_turn - initial
at_turn).

Prints one JSON line (see the "roofline" and "cpu_baseline" fields):
roofline achieved = algorithmic FP64 flops of the tracked particle-turns
(lattices.flops_per_particle_turn) / summed kernel time measured with HIP
events on the job's stream; cpu_baseline = the C99 oracle with OpenMP on
the host cores over a bounded sample of the same workload.

Beside the headline figure the line carries the other forms of the path:
"batched" (64 machine seeds in one graph of persistent launches vs 64
plain jobs),
"adjoint" (backward / forward cost of the differentiable tracker) and
"elem_by_elem" (element-by-element storage through the ring with
beam-beam and space-charge elements; an HBM-bound kernel, with its own
roofline in GB/s).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector (FMA = 2 flops)
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (KiB per dispatch; a 1 GiB device-to-device copy reads 524,308 KiB of
# FETCH_SIZE and writes 1,048,576 KiB of WRITE_SIZE, hence x2048 and x1024 B):
# a track launch of the plain 1000-turn job of the 1M-particle ring (mean of
# its 25 launches, all of them mixed launches:
# 5,670,873 / 5,064,299 KiB) and one turn of element-by-element
# tracking of the collective ring with 100k particles in the exact build
# (one 10-turn call / 10: 1,473,333 / 2,052,266,253 KiB)
TRACK_LAUNCH_HBM_BYTES = 5670873 * 2048 + 5064299 * 1024
EBE_TURN_HBM_BYTES = (1473333 * 2048 + 2052266253 * 1024) // 10


def sigma_at_start(eb):
    """rms beam size for eps_n = 2.5 um at the ring start (one-turn beta
    from the linear map of the lattice is ~236 m; computed on the host
    from the thin-lens FODO parameters used by the generator)."""
    from sixtracklib_amd import lattices
    return lattices.twiss_sigma(beta=236.0)


def build_workload(n_particles, seed=0):
    from sixtracklib_amd import lattices
    eb = lattices.lhc_like_ring(seed=seed)
    sig = sigma_at_start(eb)
    n_ang = 64
    n_amp = max(1, n_particles // n_ang)
    pb, p = lattices.polar_grid(n_amp, n_ang, sigma_x=sig, sigma_y=sig, max_sigma=20.0,
                                delta=2.7e-4)
    return eb, pb, p


def cpu_baseline(eb, pb, budget_s=10.0, sample=8192, turns_per_chunk=2):
    """Time the oracle (C99, -ffp-contract=off, OpenMP schedule(static)) on
    an evenly strided sample of the same particles; returns
    (particle-turns/s, threads, description)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from sixtracklib_amd import lattices
    from sixtracklib_amd.particles import particles_in
    p = particles_in(pb)
    n = p.num_particles
    idx = np.linspace(0, n - 1, min(sample, n)).astype(np.int64)
    sb, sp = lattices.make_particles(len(idx), x=np.array(p.x)[idx], y=np.array(p.y)[idx],
                                     delta=2.7e-4)
    threads = cpu_threads()
    at0 = np.array(sp.at_turn).copy()
    t0 = time.perf_counter()
    turn = 0
    while True:
        turn += turns_per_chunk
        oracle_lib.track_until(sb, eb, turn, threads=threads)
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    tracked = int(np.sum(np.array(sp.at_turn) - at0))
    return tracked / dt, threads, ("oracle (C99 restatement, OpenMP) on %d of the %d particles "
                                   "x %d turns of the same ring, %.1f s" % (len(idx), n, turn, dt))


def host_cores():
    """(logical CPUs of the host, physical cores of the host, CPUs this
    process may run on)."""
    logical = os.cpu_count() or 1
    phys = set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":", 1)[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":", 1)[1].strip()
                elif not line.strip():
                    if cid is not None:
                        phys.add((pid, cid))
                    pid = cid = None
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = logical
    return logical, (len(phys) or logical), allowed


def cpu_threads():
    """Threads for the CPU baseline: every CPU the process may use, up to
    OMP_NUM_THREADS when that is set (the host share given to one GPU)."""
    _, _, allowed = host_cores()
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        allowed = min(allowed, int(env))
    return max(1, allowed)


def batched_vs_sequential(n_seeds, turns):
    """Batched form: n_seeds machine seeds of the 20k-element ring (own
    multipole-error table each) with U(1000, 100000) particles per seed,
    tracked ``turns`` turns by one TrackBatch (autotuned, chunks in one HIP
    graph) and by n_seeds plain TrackJobs one after the other.  Both count
    tracked particle-turns over the wall time of their track_until calls
    (lattices and particles already on the device)."""
    from sixtracklib_amd import lattices
    from sixtracklib_amd.cobjects import CBuffer
    from sixtracklib_amd.particles import particles_in
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.trackbatch import TrackBatch
    rings, parts, counts = lattices.batch_workload(n_seeds)
    bparts = [CBuffer.frombytes(p.tobytes()) for p in parts]

    def tracked(bufs):
        return sum(int(np.sum(np.array(particles_in(b).at_turn))) for b in bufs)

    jobs = [TrackJob(eb, pb, config_str="fma=1") for eb, pb in zip(rings, parts)]
    t0 = time.perf_counter()
    for k, j in enumerate(jobs):
        j.track_until(turns)
        if (k + 1) % 8 == 0:
            progress("sequential jobs: %d of %d done, %.1f s" % (k + 1, len(jobs),
                                                                 time.perf_counter() - t0))
    seq_wall = time.perf_counter() - t0
    progress("sequential jobs: %.1f s" % seq_wall)
    for j in jobs:
        j.collect_particles()
    del jobs
    batch = TrackBatch(rings, bparts, config_str="fma=1")
    cfg = batch.autotune(trial_turns=2)
    t0 = time.perf_counter()
    batch.track_until(turns)
    b_wall = time.perf_counter() - t0
    b_kernel_ms = batch.last_kernel_time_ms
    batch.collect()
    del batch
    seq_t, b_t = tracked(parts), tracked(bparts)
    fields = ("x", "px", "y", "py", "zeta", "delta", "s", "at_element", "at_turn", "state")
    same = all(np.array_equal(np.asarray(getattr(particles_in(a), f)),
                              np.asarray(getattr(particles_in(b), f)))
               for a, b in zip(parts, bparts) for f in fields)
    return {"seeds": n_seeds, "particles": int(np.sum(counts)), "turns": turns,
            "value": b_t / b_wall, "unit": "particle-turns/s",
            "sequential_value": seq_t / seq_wall, "speedup": seq_wall / b_wall,
            "kernel_seconds": b_kernel_ms / 1e3, "tracked_particle_turns": b_t,
            "results_identical_to_sequential": bool(same and seq_t == b_t),
            "autotuned_config": cfg}


def adjoint_cost(reps=3):
    """Differentiable form: 4,096 particles, the 1,000-element sub-lattice
    (Drift, Multipole orders 1-3, Cavity), 10 turns, checkpoints every 64
    elements; loss = sum over turns of |(x, px, y, py) - target(t)|^2.
    Wall time (device synchronised) of the forward pass and of the reverse
    pass (gradients of the 6 initial coordinates of every particle and of
    every knl / ksl), best of ``reps``."""
    import torch
    from sixtracklib_amd import lattices
    from sixtracklib_amd.adjoint import AdjointTracker
    eb = lattices.adjoint_sublattice()
    _, p = lattices.gaussian_beam(4096, seed=7, sigma_x=3e-4, sigma_px=3e-6, sigma_y=3e-4,
                                  sigma_py=3e-6, sigma_zeta=0.05, sigma_delta=1e-4)
    coords = np.stack([np.array(getattr(p, f)) for f in ("x", "px", "y", "py", "zeta",
                                                          "delta")], 1)
    tgt = np.random.default_rng(11).normal(0.0, [3e-5, 3e-7, 3e-5, 3e-7], (10, 4))
    out = {}
    for fma in (False, True):
        tr = AdjointTracker(eb, 10, species=p, K=64, fma=fma)
        X0 = torch.tensor(coords, dtype=torch.float64, device="cuda")
        best_f = best_b = float("inf")
        for _ in range(reps):
            X = X0.clone().requires_grad_()
            knl = tr.knl0.clone().requires_grad_()
            ksl = tr.ksl0.clone().requires_grad_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss = tr.loss(X, knl, ksl, tgt)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            best_f, best_b = min(best_f, t1 - t0), min(best_b, t2 - t1)
        out["fma" if fma else "exact"] = {"forward_ms": best_f * 1e3, "backward_ms": best_b * 1e3,
                                          "cost_ratio": best_b / best_f}
    out.update({"particles": 4096, "elements": 1000, "turns": 10, "checkpoint_every": 64,
                "gradients": "initial 6D coordinates + all knl/ksl (%d multipoles)" % tr.n_mp})
    return out


def elem_by_elem_collective(n_particles=100000, turns=10, passes=256, reps=2, fma=False):
    """Element-by-element form: lattices.collective_ring() (20,019
    elements: the 20k ring + 2 BeamBeam6D x 5 slices + 1 BeamBeam4D + 16
    SCQGaussProfile), a 6D Gaussian beam matched at the ring start, every
    element pass of every particle stored (168 B each) for ``turns`` turns
    into a rolling store of ``passes`` passes per particle
    (ebe_store_passes; 100k x 20,019 x 10 records would be 3.4 TB).
    Achieved HBM rate = 168 B x records written / kernel time (HIP events
    on the job's stream); best of ``reps`` runs from the same initial
    state.  The exact build (fma=0) by default: bit-identical to the
    oracle, and the kernel is bound by its stores, not its arithmetic."""
    from sixtracklib_amd import lattices
    from sixtracklib_amd.particles import particles_in
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.collective_ring()
    pb, p = lattices.matched_gaussian_beam(n_particles)
    init = pb.words.copy()
    job = TrackJob(eb, pb, until_turn_elem_by_elem=turns,
                   config_str="fma=%d ebe_store_passes=%d" % (int(fma), passes))
    n_elem = eb.n_objects
    best = None
    for _ in range(reps):
        pb.words[:] = init
        job.push_particles()
        t0 = time.perf_counter()
        job.track_elem_by_elem(turns)
        wall = time.perf_counter() - t0
        kms = job.last_kernel_time_ms
        if best is None or kms < best[1]:
            best = (wall, kms)
    job.collect_particles()
    at_turn, at_elem = np.array(p.at_turn), np.array(p.at_element)
    alive = np.array(p.state) == 1
    # the same particles and turns through track_until (no records stored):
    # the two are different kernels, so in the FMA build their roundings may
    # differ (the compiler fuses each kernel's products and sums on its
    # own); reported: bit identity, the largest relative difference per field
    # (floored at the field's rms over the set) and whether loss elements,
    # turns and states are equal
    ebe_after = pb.words.copy()
    ebe_fields = {f: np.array(getattr(p, f)).copy() for f in
                  ("x", "px", "y", "py", "zeta", "delta", "psigma", "rpp", "rvv", "s",
                   "at_element", "at_turn", "state")}
    pb.words[:] = init
    job.push_particles()
    t0 = time.perf_counter()
    job.track_until(turns)
    until_wall = time.perf_counter() - t0
    job.collect_particles()
    until_tracked = int(np.sum(np.array(p.at_turn)))
    same_bits = bool(np.array_equal(pb.words, ebe_after))
    worst = 0.0
    same_index = True
    for f, a in ebe_fields.items():
        b = np.array(getattr(p, f))
        if f in ("at_element", "at_turn", "state"):
            same_index = same_index and bool(np.array_equal(a, b))
        else:
            scale = max(float(np.sqrt(np.mean(b * b))), 1e-300)
            worst = max(worst, float(np.max(np.abs(a - b) / np.maximum(np.abs(b), scale))))
    pb.words[:] = ebe_after
    # a record is stored before every element a particle enters, the one
    # where it is lost included; the first ``passes`` records of a particle
    # in a call fill its slots (168 B each), later ones overwrite the 13
    # fields that change (104 B: the 8 fixed fields already hold the
    # particle's values)
    per_particle = np.where(alive, at_turn * n_elem, at_turn * n_elem + at_elem + 1)
    records = int(np.sum(per_particle))
    first = np.minimum(per_particle, passes)
    stored_bytes = int(np.sum(first * 168 + (per_particle - first) * 104))
    wall, kms = best
    del job
    gbs = stored_bytes / (kms / 1e3) / 1e9
    return {"workload": "track_elem_by_elem, collective_ring(): %d elements (20k ring + 2 "
                        "BeamBeam6D x 5 slices + BeamBeam4D + 16 SCQGaussProfile); %d "
                        "particles, matched 6D Gaussian; %d turns; rolling store of %d "
                        "passes per particle" % (n_elem, n_particles, turns, passes),
            "build": "fma" if fma else "exact",
            "particles": n_particles, "turns": turns, "elements": n_elem,
            "store_passes": passes, "store_gb": n_particles * passes * 168 / 1e9,
            "records": records, "records_per_second": records / (kms / 1e3),
            "stored_bytes": stored_bytes, "alive_at_end": int(np.count_nonzero(alive)),
            "kernel_seconds": kms / 1e3, "wall_seconds": wall,
            "value": int(np.sum(at_turn)) / wall, "unit": "particle-turns/s",
            "track_until_value": until_tracked / until_wall,
            "track_until_vs_elem_by_elem": {"bit_identical": same_bits,
                                            "max_rel_diff": worst,
                                            "loss_turns_and_states_equal": same_index},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS,
                         "bytes_per_record": "168 for a particle's first %d records of the "
                                             "call, 104 after" % passes,
                         "traffic_per_turn": EBE_TURN_HBM_BYTES}}


def calibration_copy(nbytes):
    """One device-to-device copy of ``nbytes`` (the HIP runtime's blit
    kernel reads and writes exactly nbytes of HBM), made through PyTorch so
    that the process holds a single HIP runtime, the one the tracker
    shares."""
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def progress(msg):
    """One line per finished part of the run, on stderr (the JSON line is
    the only stdout output)."""
    print("bench.py: " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--turns", type=int, default=1000,
                    help="turns of the timed job, split evenly over the steps")
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--mode", choices=("fma", "exact"), default="fma")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact-steps", type=int, default=1,
                    help="extra timed steps with the exact (no-FMA) build, reported beside")
    ap.add_argument("--batch-seeds", type=int, default=64,
                    help="batched form: machine seeds (0 disables the batched measurement)")
    ap.add_argument("--batch-turns", type=int, default=1000,
                    help="batched form: turns tracked by the batch and by the sequential jobs")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiable-tracking cost measurement")
    ap.add_argument("--ebe-turns", type=int, default=10,
                    help="element-by-element form: turns stored (0 disables it)")
    ap.add_argument("--ebe-particles", type=int, default=100000)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device-to-device hipMemcpy of this many MiB (a copy "
                         "kernel of known traffic, to scale FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the tracker shards particles "
              "across GPUs without communication; multi-GPU runs are not part of it)",
              file=sys.stderr)
        sys.exit(2)

    from sixtracklib_amd import lattices
    from sixtracklib_amd.particles import particles_in
    from sixtracklib_amd.trackjob import TrackJob

    t_setup = time.perf_counter()
    eb, pb, _ = build_workload(args.particles)
    p = particles_in(pb)
    init_bytes = pb.words.copy()
    flops_pt, n_trans = lattices.flops_per_particle_turn(eb)
    steps = max(1, args.steps)
    # step k ends at turn targets[k]; the steps together are the whole job
    targets = [int(round(args.turns * (k + 1) / steps)) for k in range(steps)]

    def run(mode, steps, warmup):
        pb.words[:] = init_bytes
        job = TrackJob(eb, pb, config_str="fma=%d block=%d" % (mode == "fma", args.block))
        for w in range(warmup):
            job.track_until(targets[min(w, len(targets) - 1)])
        if warmup:
            pb.words[:] = init_bytes
            job.push_particles()
        at0 = np.array(p.at_turn).copy()
        kernel_ms = 0.0
        t0 = time.perf_counter()
        for k in range(steps):
            job.track_until(targets[k])
            kernel_ms += job.last_kernel_time_ms
        t1 = time.perf_counter()
        job.collect_particles()
        tracked = int(np.sum(np.array(p.at_turn) - at0))
        alive = int(np.count_nonzero(np.array(p.state) == 1))
        del job
        return tracked, t1 - t0, kernel_ms, alive

    setup_s = time.perf_counter() - t_setup
    if args.calibrate_copy_mib > 0:
        calibration_copy(args.calibrate_copy_mib << 20)
    tracked, wall, kernel_ms, alive = run(args.mode, steps, args.warmup)
    progress("plain: %d particle-turns in %.2f s" % (tracked, wall))
    value = tracked / wall
    kernel_s = kernel_ms / 1e3
    achieved_tflops = tracked * flops_pt / kernel_s / 1e12
    # HBM bytes the kernel must move per launch, for at most one 64-particle
    # tile per resident wave (32 waves per CU): per particle its list entry
    # (4 B), the 14 fields it reads (112 B: 10 coordinates, chi, at_element,
    # at_turn, state) and the 13 it writes back (104 B)
    import torch
    resident = 64 * 32 * torch.cuda.get_device_properties(0).multi_processor_count
    alg_bytes = (4 + 112 + 104) * min(args.particles, resident)

    extra = {}
    if args.exact_steps > 0:
        tr_e, wall_e, kms_e, _ = run("exact" if args.mode == "fma" else "fma",
                                     min(args.exact_steps, steps), 0)
        extra = {"other_build": "exact" if args.mode == "fma" else "fma",
                 "other_build_value": tr_e / wall_e,
                 "other_build_tflops": tr_e * flops_pt / (kms_e / 1e3) / 1e12}

    batched = None
    if args.batch_seeds > 0:
        batched = batched_vs_sequential(args.batch_seeds, args.batch_turns)
        progress("batched: %.3g particle-turns/s" % batched["value"])

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_cost()
        progress("adjoint: cost ratio %.2f" % adjoint["exact"]["cost_ratio"])

    ebe = None
    if args.ebe_turns > 0:
        ebe = elem_by_elem_collective(args.ebe_particles, args.ebe_turns)
        progress("element by element: %.0f GB/s" % ebe["roofline"]["achieved"])
        # the FMA build of the same leg, reported beside (its results are
        # within the FMA tolerance, not bit-identical)
        f = elem_by_elem_collective(args.ebe_particles, args.ebe_turns, fma=True)
        ebe["fma_build"] = {"records_per_second": f["records_per_second"],
                            "achieved_gbs": f["roofline"]["achieved"],
                            "track_until_value": f["track_until_value"]}
        progress("element by element, FMA build: %.0f GB/s" % f["roofline"]["achieved"])

    cpu = None
    if not args.no_cpu_baseline:
        v, threads, desc = cpu_baseline(eb, pb, budget_s=args.cpu_budget)
        logical, physical, allowed = host_cores()
        cpu = {"value": v, "unit": "particle-turns/s", "cores": threads, "kind": "port",
               "sample": desc, "cpu_model": cpu_model(), "host_logical_cpus": logical,
               "host_physical_cores": physical, "cpus_available_to_process": allowed,
               "per_core_value": v / threads,
               "threads_rule": "all CPUs the process may use, capped by OMP_NUM_THREADS "
                               "(the host share of one GPU)"}
        progress("cpu baseline: %.3g particle-turns/s on %d threads" % (v, threads))

    ms_per_step = wall * 1e3 / steps
    out = {
        "metric": "FP64 particle-turns/s, 20k-elem ring, 1M particles; batched & adjoint vs plain",
        "value": value,
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain track_until, lhc_like_ring(seed=0): 20000 elements "
                               "(10800 Drift, 7992 Multipole order 0-10, 400 XYShift, "
                               "400 SRotation, 400 LimitEllipse, 8 Cavity); %d particles "
                               "(polar grid 0-20 sigma x 64 angles); %d turns"
                               % (args.particles, args.turns),
                   "build": args.mode, "turns": args.turns,
                   "turns_per_step": args.turns / steps, "block": args.block,
                   "particles": args.particles, "elements": eb.n_objects,
                   "flops_per_particle_turn": flops_pt,
                   "tracked_particle_turns": tracked, "alive_at_end": alive,
                   "kernel_seconds": kernel_s, "setup_seconds": setup_s},
        "roofline": {"bound": "fp64_valu", "achieved": achieved_tflops,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP64_PEAK_TFLOPS,
                     "traffic": TRACK_LAUNCH_HBM_BYTES,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of the track launches "
                                       "(st_k_batch_track_mix_fma) on this workload, mean of the 25 "
                                       "launches of the 1000-turn job, scaled by a 1 GiB device copy "
                                       "(FETCH_SIZE reads 0.5x, WRITE_SIZE 1.0x); waves with two "
                                       "particles per lane read the particle constants and delta and "
                                       "write delta and psigma at each Cavity (registers hold only the "
                                       "Drift/Multipole state): 16.8 GB per launch at 58 GB/s",
                     "algorithmic_hbm_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "elem_by_elem": ebe,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

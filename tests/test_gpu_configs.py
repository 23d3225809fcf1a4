"""Parity at the sizes of the benchmark configurations.

* C3, the 20,019-element collective ring (20k ring + 2 BeamBeam6D x 5
  slices + BeamBeam4D + 16 SCQGaussProfile) in full: one turn of
  track_until and of track_elem_by_elem into the reference-size store,
  exact build bit for bit against the oracle; the FMA build within the
  bound measured on this lattice.
* C4, 64 machine seeds of the 20k-element ring with U(1,000, 100,000)
  particles each (3,133,879 in total): every seed of one TrackBatch equal
  bit for bit to a plain TrackJob on that seed, and a strided sample of
  every seed equal to the oracle.
* C2 loss turns in the FMA build: particles straddling the dynamic
  aperture of the 20k-element ring tracked 1000 turns in both builds; on
  regular orbits (whose exact-build loss turn does not move when the
  initial coordinates move by one ulp) the loss element, turn and state
  are those of the exact build.
* C5 at its size: 4,096 particles, 10 turns, the 1,000-element sub-lattice:
  coordinate and coefficient gradients against central differences, and
  the 10-turn 6x6 Jacobian from the adjoint against finite differences of
  the forward tracker.

Pattern: tests/sixtracklib/cuda/track/test_track_job_track_until_c99.cpp
(GPU result against the CPU tracker on the same inputs).
"""
import math

import numpy as np
import pytest

import oracle_lib
from helpers import particles_dict, assert_particles_equal
from sixtracklib_amd import lattices
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import particles_in, Particles, REAL_FIELDS

pytestmark = pytest.mark.gpu


def _copy(cb):
    return CBuffer.frombytes(cb.tobytes())


def _pdict(cb, idx=0):
    return particles_dict(particles_in(cb, idx))


def _rel_dev(got, want, fields=REAL_FIELDS):
    """Largest |a - b| / max(|b|, rms(b)) per field."""
    out = {}
    for f in fields:
        a, b = np.asarray(got[f]), np.asarray(want[f])
        scale = max(float(np.sqrt(np.mean(b * b))), 1e-300)
        out[f] = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), scale)))
    return out


# ------------------------------------------------------------------- C3

@pytest.fixture(scope="module")
def collective():
    return lattices.collective_ring()


def test_collective_ring_full_track_until_exact(collective):
    """One turn of 512 matched-beam particles through the full C3 lattice:
    exact build bit-identical to the oracle in every field."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = collective
    assert eb.n_objects == 20019
    pb, _ = lattices.matched_gaussian_beam(512, seed=21)
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 1, threads=16) == 0
    job = TrackJob(eb, pb, config_str="fma=0")
    assert job.uses_full_kernel
    job.track_until(1).collect()
    want = _pdict(ob)
    assert np.count_nonzero(want["at_turn"] == 1) == 512
    assert_particles_equal(_pdict(pb), want, what="collective ring, exact")


def test_collective_ring_full_elem_by_elem_reference_store(collective):
    """One turn element by element through the full C3 lattice into the
    reference's own store (particles x 20,019 elements x 1 turn = 2.6 M
    records): every record and the final particles bit-identical to the
    oracle's."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = collective
    E = eb.n_objects
    n = 128
    pb, _ = lattices.matched_gaussian_beam(n, seed=22)
    ob = _copy(pb)
    job = TrackJob(eb, pb, until_turn_elem_by_elem=1, config_str="fma=0")
    off = job.elem_by_elem_output_offset
    job.track_elem_by_elem(1).collect()
    got = _pdict(job.output, off)
    assert len(got["x"]) == n * E
    ocb = CBuffer(max_slots=22 * n * E + 64)
    Particles(cbuffer=ocb, num_particles=n * E)
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=n, num_elements_to_store=E,
                               num_turns_to_store=1, min_particle_id=0, min_element_id=0,
                               min_turn=0, max_particle_id=n - 1, max_element_id=E - 1,
                               max_turn=0, is_rolling=1, out_store_addr=ocb.get_object_address(0))
    assert oracle_lib.track_elem_by_elem(ob, eb, cfg, 1) == 0
    assert_particles_equal(_pdict(pb), _pdict(ob), what="particles")
    assert_particles_equal(got, _pdict(ocb), what="C3 element-by-element store")


# Measured on this lattice (4,096 matched particles, one turn): the FMA
# build's largest difference from the exact build, relative to each
# field's rms, is 1.35e-13 (py); x, px, y 5-6e-14; zeta 1.2e-16; delta,
# psigma, rpp, rvv identical.  The beam-beam and space-charge kicks
# amplify the roundings of the 18,000 fused drift and multipole updates
# that precede them, so this lattice is held to 2e-13, the 20,000-element
# ring without them to 1e-13.
C3_FMA_REL = 2e-13


def test_collective_ring_full_fma_one_turn(collective):
    from sixtracklib_amd.trackjob import TrackJob
    eb = collective
    pb, _ = lattices.matched_gaussian_beam(1024, seed=23)
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 1, threads=16) == 0
    job = TrackJob(eb, pb, config_str="fma=1")
    job.track_until(1).collect()
    got, want = _pdict(pb), _pdict(ob)
    for f in ("particle_id", "at_element", "at_turn", "state"):
        assert np.array_equal(got[f], want[f]), f
    dev = _rel_dev(got, want)
    worst = max(dev.values())
    assert worst <= C3_FMA_REL, dev
    for f in ("delta", "psigma", "rpp", "rvv"):
        assert dev[f] == 0.0, dev


# ------------------------------------------------------------------- C4

def test_batch_c4_full_size_against_plain_jobs_and_oracle():
    """The C4 workload (64 seeds, 3.13 M particles): 3 turns by one
    TrackBatch (exact build) and by 64 plain TrackJobs; every seed's
    particles bit-identical between the two, and every 97th particle of
    every seed bit-identical to the oracle."""
    from sixtracklib_amd.trackbatch import TrackBatch
    from sixtracklib_amd.trackjob import TrackJob
    rings, parts, counts = lattices.batch_workload(64)
    assert int(np.sum(counts)) == 3133879
    turns = 3
    bparts = [_copy(p) for p in parts]
    samples = []
    for pb in parts:
        p = particles_in(pb)
        idx = np.arange(0, p.num_particles, 97)
        sb, _ = lattices.make_particles(len(idx), x=np.array(p.x)[idx], y=np.array(p.y)[idx],
                                        delta=np.array(p.delta)[idx],
                                        particle_id=np.array(p.particle_id)[idx])
        samples.append((idx, sb))
    batch = TrackBatch(rings, bparts, config_str="fma=0")
    assert batch.track_until(turns) == 0
    batch.collect()
    del batch
    for s, (eb, pb) in enumerate(zip(rings, parts)):
        job = TrackJob(eb, pb, config_str="fma=0")
        job.track_until(turns).collect()
        del job
        want = _pdict(pb)
        assert_particles_equal(_pdict(bparts[s]), want, what="seed %d batch vs plain" % s)
        idx, sb = samples[s]
        assert oracle_lib.track_until(sb, eb, turns, threads=16) == 0
        sub = {f: v[idx] for f, v in want.items()}
        assert_particles_equal(sub, _pdict(sb), what="seed %d sample vs oracle" % s)


# ------------------------------------------------------- C2 loss turns

def test_fma_loss_turns_equal_exact_on_regular_orbits():
    """4,096 particles at 6-12 sigma (the dynamic aperture of the
    20,000-element ring lies there) x 64 angles, 1000 turns, delta
    2.7e-4.  Regular orbits: the exact build gives the same loss element,
    turn and state when x and y start one ulp larger.  On those the FMA
    build's loss element, turn and state equal the exact build's."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.lhc_like_ring(seed=0)
    sig = lattices.twiss_sigma(beta=236.0)
    r = np.linspace(6.0, 12.0, 64)
    th = (np.arange(64) + 0.5) * (0.5 * np.pi / 64)
    rr, tt = np.meshgrid(r, th, indexing="ij")
    x = (rr * np.cos(tt) * sig).ravel()
    y = (rr * np.sin(tt) * sig).ravel()
    n = x.size

    def run(xs, ys, fma):
        pb, _ = lattices.make_particles(n, x=xs, y=ys, delta=2.7e-4)
        TrackJob(eb, pb, config_str="fma=%d" % int(fma)).track_until(1000).collect()
        d = _pdict(pb)
        return d["at_turn"], d["at_element"], d["state"]

    ref = run(x, y, False)
    pert = run(np.nextafter(x, np.inf), np.nextafter(y, np.inf), False)
    fma = run(x, y, True)
    lost = ref[2] == 0
    regular = (ref[0] == pert[0]) & (ref[1] == pert[1]) & (ref[2] == pert[2])
    n_lost, n_reg = int(np.count_nonzero(lost)), int(np.count_nonzero(regular))
    # the set straddles the aperture: particles lost at many turns, most
    # of them on regular orbits
    assert 0.2 * n < n_lost < 0.9 * n, n_lost
    assert len(np.unique(ref[0][lost])) > 50
    assert n_reg > 0.8 * n, n_reg
    bad = regular & ((fma[0] != ref[0]) | (fma[1] != ref[1]) | (fma[2] != ref[2]))
    assert not np.any(bad), ("FMA loss differs on %d regular orbits of %d (%d lost)"
                             % (int(np.count_nonzero(bad)), n_reg, n_lost))


# ------------------------------------------------------------------- C5

C5_FIELDS = ("x", "px", "y", "py", "zeta", "delta")
C5_SIGMA = np.array([3e-4, 3e-6, 3e-4, 3e-6, 0.05, 1e-4])


def _c5():
    from sixtracklib_amd.adjoint import AdjointTracker
    eb = lattices.adjoint_sublattice()
    _, p = lattices.gaussian_beam(4096, seed=7, sigma_x=3e-4, sigma_px=3e-6, sigma_y=3e-4,
                                  sigma_py=3e-6, sigma_zeta=0.05, sigma_delta=1e-4)
    coords = np.stack([np.array(getattr(p, f)) for f in C5_FIELDS], 1)
    tgt = np.random.default_rng(11).normal(0.0, [3e-5, 3e-7, 3e-5, 3e-7], (10, 4))
    return eb, p, coords, tgt, AdjointTracker


def _dev(a):
    torch = pytest.importorskip("torch")
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def test_c5_gradients_full_size_match_central_differences():
    """4,096 particles, 10 turns, K = 64: d loss / d (initial coordinates)
    of 6 particles spread over the beam, and d loss / d knl, ksl of 10
    coefficients (every order), against Richardson-extrapolated central
    differences (steps 1e-4 of the beam size / coefficient scale): within
    1e-8 relative (coordinates: floored at 1e-8 of the particle's largest
    component; coefficients: plus the differences' own noise,
    8e-16 |loss| / h)."""
    eb, p, coords, tgt, AdjointTracker = _c5()
    tr = AdjointTracker(eb, 10, species=p, K=64)
    X = _dev(coords).requires_grad_()
    knl = tr.knl0.clone().requires_grad_()
    ksl = tr.ksl0.clone().requires_grad_()
    loss = tr.loss(X, knl, ksl, tgt)
    loss.backward()
    L = float(loss.detach())
    g_adj = X.grad.cpu().numpy()
    pick = [0, 700, 1501, 2222, 3333, 4095]
    steps = (1.0, -1.0, 0.5, -0.5)
    pert = []
    for i in pick:
        for c in range(6):
            for st in steps:
                v = coords[i].copy()
                v[c] += st * 1e-4 * C5_SIGMA[c]
                pert.append(v)
    lp = tr.per_particle_loss(_dev(np.array(pert)), target=tgt).cpu().numpy().reshape(-1, 4)
    h = np.tile(1e-4 * C5_SIGMA, len(pick))
    d1 = (lp[:, 0] - lp[:, 1]) / (2 * h)
    d2 = (lp[:, 2] - lp[:, 3]) / h
    g_fd = ((4 * d2 - d1) / 3).reshape(len(pick), 6)
    ga = g_adj[pick]
    scale = np.max(np.abs(g_fd), axis=1, keepdims=True)
    err = np.abs(ga - g_fd) / (np.abs(g_fd) + scale)
    assert np.max(err) < 1e-8, (np.max(err), ga, g_fd)

    gk, gs = knl.grad.cpu().numpy(), ksl.grad.cpu().numpy()
    mask = tr.mask.cpu().numpy()
    k0, s0 = tr.knl0.cpu().numpy(), tr.ksl0.cpu().numpy()
    big = max(np.max(np.abs(gk)), np.max(np.abs(gs)))
    valid = np.argwhere(mask)
    rng = np.random.default_rng(5)
    sample = valid[rng.choice(len(valid), 5, replace=False)]
    for (m, nn) in sample:
        err_scale = 1e-4 * 0.0207 * math.factorial(int(nn)) * 0.017 ** (1.0 - nn)
        for which, base, g in (("knl", k0, gk), ("ksl", s0, gs)):
            hh = 1e-4 * max(abs(base[m, nn]), err_scale)

            def f(step):
                kk, ss = k0.copy(), s0.copy()
                (kk if which == "knl" else ss)[m, nn] += step
                return float(tr.per_particle_loss(_dev(coords), _dev(kk), _dev(ss), tgt).sum())
            fd = (4 * (f(hh / 2) - f(-hh / 2)) / hh - (f(hh) - f(-hh)) / (2 * hh)) / 3
            noise = 8 * 2.2e-16 * abs(L) / hh
            assert abs(g[m, nn] - fd) <= 1e-8 * abs(fd) + noise + 1e-12 * big, \
                (which, m, nn, g[m, nn], fd, noise)


def test_c5_ten_turn_jacobian_matches_forward_differences():
    """The 10-turn 6x6 Jacobian of one particle of the C5 beam: from the
    adjoint, J^T e_c = (g(T) - g(T + d e_c)) / (2 d) with unit weights and
    the target moved in the last turn only (the loss terms of turns 0-8 do
    not depend on it); from TrackJob (exact build) by Richardson-
    extrapolated central differences of the 10-turn state.  Steps: 1e-4
    sigma_c for the x, px, y, py rows; 0.3 sigma_c for the zeta and delta
    rows, whose differences are limited by rounding, not truncation: each
    Cavity pass leaves delta with ~1e-16 absolute error (a square root near
    1, minus 1) and each drift zeta with ~1e-16 * L, so steps of 1e-2
    sigma_px give 2e-8 errors in d delta / d px0, while 0.3 sigma keeps them
    below 1e-9 (oracle differences at steps 1e-4 ... 1 sigma).  Entries
    within 1e-8 of the largest entry of their row; the zeta row within 3e-7
    (4.6e-8 measured).
    """
    from sixtracklib_amd.trackjob import TrackJob
    eb, p, coords, _, AdjointTracker = _c5()
    turns = 10
    tr = AdjointTracker(eb, turns, species=p, weights=(1.0,) * 6, K=64)
    x0 = coords[1234]

    def grad_for(last_target):
        tg = np.zeros((turns, 6))
        tg[-1] = last_target
        X = _dev(x0[None, :]).requires_grad_()
        tr.loss(X, target=tg).backward()
        return X.grad.cpu().numpy()[0]

    g0 = grad_for(np.zeros(6))
    d = 1e-3
    J_adj = np.zeros((6, 6))
    for c in range(6):
        e = np.zeros(6)
        e[c] = d
        J_adj[c] = (g0 - grad_for(e)) / (2 * d)

    def fd_jacobian(rel):
        starts = []
        for c in range(6):
            for st in (1.0, -1.0, 0.5, -0.5):
                v = x0.copy()
                v[c] += st * rel * C5_SIGMA[c]
                starts.append(v)
        starts = np.array(starts)
        pb, _ = lattices.make_particles(len(starts),
                                        **{f: starts[:, k] for k, f in enumerate(C5_FIELDS)},
                                        p0c=float(p.p0c[0]), mass0=float(p.mass0[0]))
        TrackJob(eb, pb, config_str="fma=0").track_until(turns).collect()
        q = particles_in(pb)
        assert np.all(np.array(q.state) == 1)
        end = np.stack([np.array(getattr(q, f)) for f in C5_FIELDS], 1).reshape(6, 4, 6)
        h = (rel * C5_SIGMA)[:, None]
        d1 = (end[:, 0] - end[:, 1]) / (2 * h)
        d2 = (end[:, 2] - end[:, 3]) / h
        return ((4 * d2 - d1) / 3).T
    J_fd = np.vstack([fd_jacobian(1e-4)[:4], fd_jacobian(0.3)[4:]])
    row_scale = np.max(np.abs(J_fd), axis=1, keepdims=True)
    err = np.abs(J_adj - J_fd) / row_scale
    assert np.max(err[[0, 1, 2, 3, 5]]) < 1e-8, (err, J_adj, J_fd)
    assert np.max(err[4]) < 3e-7, (err, J_adj, J_fd)

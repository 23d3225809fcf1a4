"""Batched multi-seed tracking (TrackBatch) against the plain path and the
oracle.

The reference has no batched form; its users run one job per seed.  The
parity bar is therefore: every seed of a batch ends bit-identical to a
plain TrackJob on that seed alone (same build), and, for the exact build,
to the CPU oracle -- with uneven particle counts (not multiples of 64, one
seed smaller than a wave), balanced launches (the default) and
graph-captured chunks, XCD-grouped and single queues, one and two
particles per lane, and after the autotuner.
"""
import numpy as np
import pytest

import oracle_lib
from helpers import assert_particles_equal
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd import lattices

pytestmark = pytest.mark.gpu

COUNTS = [37, 500, 1000, 64, 129, 2000]


def _copy(cb):
    return CBuffer.frombytes(cb.tobytes())


def _pdict(cb):
    from sixtracklib_amd.particles import particles_in
    from helpers import particles_dict
    return particles_dict(particles_in(cb))


def _workload(n_cells=8, amp=30.0):
    """6 seeds of a 400-element ring (8 cells of the 20k-element ring's
    structure, one cavity), error tables 0..5, polar-grid particles up to
    ``amp`` sigma (some lost within a few turns)."""
    rings = lattices.lhc_like_ring_family(len(COUNTS), seed=0, n_cells=n_cells, n_cavities=1)
    sig = lattices.twiss_sigma(beta=236.0)
    parts = []
    for n in COUNTS:
        n_amp = (n + 63) // 64
        pb, _ = lattices.polar_grid(n_amp, 64, sigma_x=sig * 2, sigma_y=sig * 2, max_sigma=amp)
        from sixtracklib_amd.particles import particles_in
        p = particles_in(pb)
        x = np.array(p.x)[:n]
        y = np.array(p.y)[:n]
        sb, _ = lattices.make_particles(n, x=x, y=y, delta=2.7e-4)
        parts.append(sb)
    return rings, parts


def _plain(rings, parts, turns, fma):
    from sixtracklib_amd.trackjob import TrackJob
    out = []
    for eb, pb in zip(rings, parts):
        job = TrackJob(eb, pb, config_str="fma=%d" % fma)
        job.track_until(turns)
        job.collect()
        out.append(_pdict(pb))
    return out


@pytest.mark.parametrize("cfg", ["fma=0", "fma=0 balance=0", "fma=0 balance=0 queues=1 graph=0",
                                 "fma=0 balance=0 turns_per_launch=2",
                                 "fma=0 balance=0 block=64 waves_per_cu=16 turns_per_launch=3",
                                 "fma=0 balance=0 ppl=2 turns_per_launch=2", "fma=0 ppl=2",
                                 "fma=0 balance=1 waves_per_cu=1 queues=8",
                                 "fma=0 balance=0 staging=lds",
                                 "fma=0 balance=0 staging=lds turns_per_launch=2 queues=1",
                                 "fma=0 staging=lds balance=1 waves_per_cu=4"])
def test_batch_matches_oracle_exact(cfg):
    from sixtracklib_amd.trackbatch import TrackBatch
    rings, parts = _workload()
    want = []
    for eb, pb in zip(rings, parts):
        ob = _copy(pb)
        oracle_lib.track_until(ob, eb, 6)
        want.append(_pdict(ob))
    batch = TrackBatch(rings, parts, config_str=cfg)
    assert batch.num_seeds == len(COUNTS) and batch.total_particles == sum(COUNTS)
    for t in (2, 6):                                # two calls: state carried on device
        batch.track_until(t)
    batch.collect()
    lost = 0
    for s, pb in enumerate(parts):
        got = _pdict(pb)
        lost += int(np.count_nonzero(got["state"] == 0))
        assert_particles_equal(got, want[s], what="seed %d (%s)" % (s, cfg))
    assert 0 < lost < sum(COUNTS)


@pytest.mark.parametrize("cfg", ["fma=1 turns_per_launch=2", "fma=1 staging=lds turns_per_launch=2",
                                 "fma=1 ppl=2", "fma=1 ppl=2 balance=0 turns_per_launch=3"])
def test_batch_fma_matches_plain_fma_bitwise(cfg):
    """FMA build: the batch (scalar-load or LDS-staged lattice, one or two
    particles per lane, balanced or chunked launches) equals the plain FMA
    job bit for bit."""
    from sixtracklib_amd.trackbatch import TrackBatch
    rings, parts = _workload()
    plain_in = [_copy(pb) for pb in parts]
    want = _plain(rings, plain_in, 5, fma=1)
    batch = TrackBatch(rings, parts, config_str=cfg)
    batch.track_until(5)
    batch.collect()
    for s, pb in enumerate(parts):
        assert_particles_equal(_pdict(pb), want[s], what="seed %d %s" % (s, cfg))


def test_batch_autotune_keeps_state_and_results():
    from sixtracklib_amd.trackbatch import TrackBatch
    rings, parts = _workload()
    plain_in = [_copy(pb) for pb in parts]
    want = _plain(rings, plain_in, 4, fma=0)
    batch = TrackBatch(rings, parts, config_str="fma=0")
    cfg = batch.autotune(trial_turns=2)
    assert "block=" in cfg and "queues=" in cfg and "staging=" in cfg and "ppl=" in cfg
    batch.track_until(4)
    batch.collect()
    for s, pb in enumerate(parts):
        assert_particles_equal(_pdict(pb), want[s], what="seed %d after autotune" % s)


def test_batch_push_restarts_from_host_state():
    from sixtracklib_amd.trackbatch import TrackBatch
    rings, parts = _workload()
    init = [_copy(pb) for pb in parts]
    want = _plain(rings, [_copy(pb) for pb in parts], 3, fma=0)
    batch = TrackBatch(rings, parts, config_str="fma=0")
    batch.track_until(7)
    for pb, ib in zip(parts, init):                 # reset the host buffers, push
        pb.words[:] = ib.words
        pb._remap()
    batch.push()
    batch.track_until(3)
    batch.collect()
    for s, pb in enumerate(parts):
        assert_particles_equal(_pdict(pb), want[s], what="seed %d after push" % s)

"""GPU parity of the balanced scheduling of track_until (the default with
compact=1): every launch gives each resident wave one tile, the particles
furthest behind first, until every particle has reached the requested turn.

Each particle's map sequence does not depend on which wave tracks it or in
how many launches, so the results are bit-identical to the oracle (exact
build) and to the persistent-chunk form (balance=0) in either build:

* more tiles than resident waves, forced at small sizes with
  ``waves_per_cu=1`` (256 waves on 256 CUs), with particles starting at
  different turns, some already lost, and two particle sets of one job;
* the same for two particles per lane, and for a lattice with the maps
  that run out of line (beam-beam, space charge, RF multipole);
* the full-occupancy regime (9,375 tiles for 8,192 resident waves) against
  balance=0, and a sample of its particles against the oracle;
* the issue priority of the waves (wave_prio) changes only timing: with it
  off the results are the same.
"""
import numpy as np
import pytest

import oracle_lib
from helpers import particles_dict, assert_particles_equal, particles_from_arrays
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import particles_in, ALL_FIELDS
from sixtracklib_amd import lattices

pytestmark = pytest.mark.gpu


def _copy(cb):
    return CBuffer.frombytes(cb.tobytes())


def _two_sets(n0, n1, seed, sigma=4e-4):
    """Two Gaussian particle sets in one buffer; at_turn 0, 2 or 5 and
    every 97th particle already lost."""
    cb = CBuffer(max_slots=22 * (n0 + n1) + 256, max_objects=8, max_pointers=64)
    rng = np.random.default_rng(seed)
    for n in (n0, n1):
        lattices.make_particles(
            n, x=rng.normal(0, sigma, n), px=rng.normal(0, 4e-6, n),
            y=rng.normal(0, sigma, n), py=rng.normal(0, 4e-6, n),
            zeta=rng.normal(0, 0.05, n), delta=rng.normal(0, 1e-4, n),
            at_turn=rng.choice([0, 2, 5], n), cbuffer=cb,
            state=np.where(np.arange(n) % 97 == 3, 0, 1))
    return cb


def _run(eb, pb, cfg, steps):
    from sixtracklib_amd.trackjob import TrackJob
    job = TrackJob(eb, pb, config_str=cfg, particle_set_indices=[0, 1])
    for t in steps:
        assert job.track_until(t).last_track_status == 0
    job.collect()
    return [particles_dict(particles_in(pb, k)) for k in range(2)]


@pytest.mark.parametrize("ppl", [1, 2])
def test_more_tiles_than_waves_mixed_turns_match_oracle(ppl):
    eb = lattices.fodo_ring(n_cells=10)
    pb = _two_sets(40000, 6001, seed=11 + ppl, sigma=6e-3)
    ob, qb = _copy(pb), _copy(pb)
    for k in range(2):
        oracle_lib.track_until(ob, eb, 15, index=k)
    want = [particles_dict(particles_in(ob, k)) for k in range(2)]
    rb = _copy(pb)
    got = _run(eb, pb, "fma=0 waves_per_cu=1 ppl=%d" % ppl, (4, 9, 15))
    ref = _run(eb, qb, "fma=0 waves_per_cu=1 balance=0 ppl=%d" % ppl, (4, 9, 15))
    nop = _run(eb, rb, "fma=0 waves_per_cu=1 wave_prio=0 ppl=%d" % ppl, (4, 9, 15))
    lost = np.count_nonzero(want[0]["state"] == 0)
    assert 500 < lost < 39000, lost
    assert np.all(want[0]["at_turn"][want[0]["state"] == 1] == 15)
    for k in range(2):
        assert_particles_equal(got[k], want[k], what="balanced, set %d" % k)
        assert_particles_equal(ref[k], want[k], what="chunked, set %d" % k)
        assert_particles_equal(nop[k], want[k], what="balanced without priority, set %d" % k)


def test_more_tiles_than_waves_out_of_line_maps_match_oracle():
    """kitchen_sink() holds every element type the tracker maps, among them
    beam-beam, space charge and an RF multipole, which run out of line (the
    heavy loop); about 60 % of the particles are lost within 8 turns."""
    eb = lattices.kitchen_sink(1)
    pb = _two_sets(18000, 3000, seed=7, sigma=2e-3)
    ob = _copy(pb)
    for k in range(2):
        oracle_lib.track_until(ob, eb, 8, index=k)
    got = _run(eb, pb, "fma=0 waves_per_cu=1", (3, 8))
    for k in range(2):
        assert_particles_equal(got[k], particles_dict(particles_in(ob, k)),
                               what="balanced heavy, set %d" % k)


@pytest.mark.parametrize("fma", [0, 1])
def test_full_occupancy_balanced_equals_chunked(fma):
    """600,000 particles (9,375 tiles) at the default occupancy, 40 turns in
    three requests: balanced and persistent-chunk results are identical,
    and (exact build) every 499th particle equals the oracle's."""
    eb = lattices.fodo_ring()
    sig = lattices.twiss_sigma(beta=20.0)
    pb, _ = lattices.polar_grid(9375, 64, sigma_x=sig, sigma_y=sig, max_sigma=12.0,
                                delta=1e-4)
    qb = _copy(pb)
    sample = {f: np.array(getattr(particles_in(pb), f))[::499] for f in ALL_FIELDS}
    got = _run1(eb, pb, "fma=%d" % fma)
    ref = _run1(eb, qb, "fma=%d balance=0" % fma)
    lost = np.count_nonzero(got["state"] == 0)
    assert 10000 < lost < 590000, lost
    assert_particles_equal(got, ref, what="balanced vs chunked")
    if not fma:
        sb, _ = particles_from_arrays(sample)
        oracle_lib.track_until(sb, eb, 40)
        sub = {f: v[::499] for f, v in got.items()}
        assert_particles_equal(sub, particles_dict(particles_in(sb)), what="sample vs oracle")


def _run1(eb, pb, cfg):
    from sixtracklib_amd.trackjob import TrackJob
    job = TrackJob(eb, pb, config_str=cfg)
    for t in (7, 23, 40):
        assert job.track_until(t).last_track_status == 0
    job.collect()
    return particles_dict(particles_in(pb))

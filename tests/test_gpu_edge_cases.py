"""GPU edge cases of the tracking loops, each against the oracle on the
same inputs (exact build, bit for bit) or against the reference's stated
semantics:

* an empty lattice is a failure (status -1) that leaves particles as they
  were (track/track.h:610-670: be_begin == be_end);
* a request for a turn already reached is a no-op with status 0, and so is
  a set whose particles are all lost;
* at_element is an offset, not a start position: tracking always begins at
  element 0 and at_element returns to its initial value at turn end
  (track.h:625-660);
* non-finite coordinates are lost at the first drift's global aperture;
* an unknown element type loses the particle with status -1
  (track.h:577-581), a disabled map (TriCub) likewise;
* particle counts of 1, 63, 65 and 129 (partial waves and tiles);
* several particle sets in one job (particle_set_indices) are tracked, a
  set not named is left untouched.
"""
import numpy as np
import pytest

import oracle_lib
from helpers import particles_dict, assert_particles_equal
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import particles_in, Particles
from sixtracklib_amd import lattices, beam_elements as be

pytestmark = pytest.mark.gpu


def _copy(cb):
    return CBuffer.frombytes(cb.tobytes())


def _job(eb, pb, **kw):
    from sixtracklib_amd.trackjob import TrackJob
    return TrackJob(eb, pb, config_str="fma=0", **kw)


def test_empty_lattice_fails_and_leaves_particles():
    from sixtracklib_amd import stcommon as st
    eb = CBuffer()
    pb, _ = lattices.gaussian_beam(100, seed=1)
    before = particles_dict(particles_in(pb))
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 3) != 0
    job = _job(eb, pb)
    assert st.st_TrackJob_track_until(job.ptr_st_track_job, 3) == -1
    job.collect()
    assert_particles_equal(particles_dict(particles_in(pb)), before, what="empty lattice")
    assert_particles_equal(particles_dict(particles_in(pb)), particles_dict(particles_in(ob)),
                           what="empty lattice vs oracle")


def test_reached_turns_and_lost_sets_are_no_ops():
    eb = lattices.fodo_ring(n_cells=5)
    pb, _ = lattices.gaussian_beam(300, seed=2, sigma_x=3e-4, sigma_y=3e-4)
    job = _job(eb, pb)
    assert job.track_until(4).last_track_status == 0
    job.collect()
    after4 = particles_dict(particles_in(pb))
    for t in (4, 3, 0):
        assert job.track_until(t).last_track_status == 0
    job.collect()
    assert_particles_equal(particles_dict(particles_in(pb)), after4, what="repeated requests")
    lb, _ = lattices.gaussian_beam(70, seed=3, state=0)
    before = particles_dict(particles_in(lb))
    job = _job(eb, lb)
    assert job.track_until(10).last_track_status == 0
    job.collect()
    assert_particles_equal(particles_dict(particles_in(lb)), before, what="all lost")


@pytest.mark.parametrize("n", [1, 63, 65, 129])
def test_at_element_offset_and_partial_waves(n):
    """Particles starting with at_element = 7 and at_turn = 2: the loss
    element is reported as 7 + index and at_element returns to 7."""
    eb = lattices.fodo_ring(n_cells=10)
    x = np.linspace(1e-4, 4e-2, n)
    pb, _ = lattices.make_particles(n, x=x, y=0.5 * x, at_element=7, at_turn=2)
    ob = _copy(pb)
    oracle_lib.track_until(ob, eb, 12)
    job = _job(eb, pb)
    assert job.track_until(12).last_track_status == 0
    job.collect()
    got, want = particles_dict(particles_in(pb)), particles_dict(particles_in(ob))
    assert_particles_equal(got, want, what="offset start")
    alive = want["state"] == 1
    assert np.all(want["at_element"][alive] == 7) and np.all(want["at_turn"][alive] == 12)
    if n > 1:                # lost at the LimitRect, element 41 of 42
        assert np.any(~alive) and np.all(want["at_element"][~alive] == 7 + 41)


def test_non_finite_coordinates_are_lost_at_the_first_drift():
    eb = lattices.fodo_ring(n_cells=3)
    x = np.array([0.0, np.nan, np.inf, -np.inf, 1e-3, 0.0])
    y = np.array([np.nan, 0.0, 0.0, 0.0, 1e-3, 2.0])
    pb, _ = lattices.make_particles(6, x=x, y=y)
    ob = _copy(pb)
    oracle_lib.track_until(ob, eb, 2)
    job = _job(eb, pb)
    assert job.track_until(2).last_track_status == 0
    job.collect()
    got = particles_dict(particles_in(pb))
    assert_particles_equal(got, particles_dict(particles_in(ob)), what="non-finite")
    assert list(got["state"]) == [0, 0, 0, 0, 1, 0]
    assert list(got["at_element"][:4]) == [0, 0, 0, 0]


@pytest.mark.parametrize("type_id", [999, 261])
def test_unknown_element_loses_particles_with_failure(type_id):
    """An element of an unknown type id (999), or of the TriCub map that
    the reference's default build disables (261), stops every particle
    there with state 0 and a -1 status, as the oracle does."""
    from sixtracklib_amd import stcommon as st
    eb = CBuffer()
    be.Drift(cbuffer=eb, length=1.0)
    eb.add_object(type_id, np.zeros(4, dtype=np.uint64))
    be.Drift(cbuffer=eb, length=1.0)
    pb, _ = lattices.gaussian_beam(80, seed=4)
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 2) != 0
    job = _job(eb, pb)
    assert st.st_TrackJob_track_until(job.ptr_st_track_job, 2) == -1
    job.collect()
    got = particles_dict(particles_in(pb))
    assert_particles_equal(got, particles_dict(particles_in(ob)), what="unknown element")
    assert np.all(got["state"] == 0) and np.all(got["at_element"] == 1)


def test_selected_particle_sets_only():
    eb = lattices.fodo_ring(n_cells=6)
    cb = CBuffer(max_slots=22 * 700 + 256, max_objects=8, max_pointers=80)
    sizes = (200, 150, 300)
    for k, n in enumerate(sizes):
        lattices.make_particles(n, x=np.linspace(1e-4, 3e-3, n) * (k + 1), cbuffer=cb)
    ob = _copy(cb)
    untouched = particles_dict(particles_in(cb, 1))
    for k in (0, 2):
        oracle_lib.track_until(ob, eb, 5, index=k)
    job = _job(eb, cb, particle_set_indices=[0, 2])
    assert job.track_until(5).last_track_status == 0
    job.collect()
    for k in (0, 2):
        assert_particles_equal(particles_dict(particles_in(cb, k)),
                               particles_dict(particles_in(ob, k)), what="set %d" % k)
    assert_particles_equal(particles_dict(particles_in(cb, 1)), untouched, what="set 1 untouched")

"""NS(TrackJob_track_until_noalloc) against NS(TrackJob_track_until).

The no-allocation entry point runs the same scheduling and kernels with its
scratch arrays in the library's static device array.  Bar: bit-identical
particle sets (exact and FMA builds) on two calls in a row, the second with
a smaller input, so that the second call starts from the first one's
leftovers in that array; and an error, not a fault, for an input beyond the
array's limit.
"""
import numpy as np
import pytest

from helpers import particles_dict, assert_particles_equal
from sixtracklib_amd import lattices
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import particles_in

pytestmark = pytest.mark.gpu


def _copy(cb):
    return CBuffer.frombytes(cb.tobytes())


def _beam(n, seed):
    sig = lattices.twiss_sigma(beta=236.0)
    n_amp = (n + 63) // 64
    pb, p = lattices.polar_grid(n_amp, 64, sigma_x=2 * sig, sigma_y=2 * sig, max_sigma=40.0,
                                delta=2.7e-4)
    x = np.array(p.x)[:n]
    y = np.array(p.y)[:n]
    rng = np.random.default_rng(seed)
    sb, _ = lattices.make_particles(n, x=x * (1 + 0.01 * rng.standard_normal(n)), y=y,
                                    delta=2.7e-4)
    return sb


def _track(eb, pb, turns, fma, noalloc, extra=""):
    from sixtracklib_amd import stcommon as st
    from sixtracklib_amd.trackjob import TrackJob
    job = TrackJob(eb, pb, config_str="fma=%d %s" % (int(fma), extra))
    for t in turns:
        f = st.st_TrackJob_track_until_noalloc if noalloc else st.st_TrackJob_track_until
        status = f(job.ptr_st_track_job, int(t))
        assert status == 0, (status, st.last_error())
    job.collect_particles()
    return particles_dict(particles_in(pb))


@pytest.mark.parametrize("fma", [False, True])
def test_noalloc_matches_track_until_twice_second_smaller(fma):
    eb = lattices.lhc_like_ring(seed=1, n_cells=8, tune=8.31 * 8 / 40)
    for n, seed, turns in ((5000, 0, (3, 8)), (700, 1, (5,))):
        pb = _beam(n, seed)
        ref_b, na_b = _copy(pb), _copy(pb)
        ref = _track(eb, ref_b, turns, fma, noalloc=False)
        got = _track(eb, na_b, turns, fma, noalloc=True)
        lost = np.count_nonzero(ref["state"] == 0)
        assert 0 < lost < n, "the input should lose some particles (%d of %d)" % (lost, n)
        assert_particles_equal(got, ref, what="noalloc %d particles" % n)


def test_noalloc_heavy_lattice_and_chunked_form():
    """Beam-beam / space-charge lattice (the out-of-line maps) and the
    chunked scheduling (balance=0), both against track_until."""
    eb = lattices.collective_ring(n_cells=8, n_sc=4)
    pb, _ = lattices.matched_gaussian_beam(1500, seed=4)
    for extra in ("", "balance=0 turns_per_launch=2"):
        ref_b, na_b = _copy(pb), _copy(pb)
        ref = _track(eb, ref_b, (2, 5), False, noalloc=False, extra=extra)
        got = _track(eb, na_b, (2, 5), False, noalloc=True, extra=extra)
        assert_particles_equal(got, ref, what="noalloc heavy %r" % extra)


def test_noalloc_rejects_more_particles_than_its_array_holds():
    from sixtracklib_amd import stcommon as st
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.fodo_ring()
    n = (1 << 21) + 64
    pb, _ = lattices.make_particles(n, x=np.full(n, 1e-4), delta=0.0)
    job = TrackJob(eb, pb, config_str="fma=0")
    assert st.st_TrackJob_track_until_noalloc(job.ptr_st_track_job, 1) == -1
    assert "at most" in st.last_error()
    job.collect_particles()
    assert np.all(np.array(particles_in(pb).at_turn) == 0)

"""GPU parity: the HIP tracker through the C-ABI / TrackJob against the CPU
oracle on the same inputs.

* exact build (fma=0): bit-identical to the oracle (stmath build) on every
  field, for every element type;
* fma build (fma=1): within 1e-13 relative (per field, floored by the
  field's scale) after one turn, identical loss turns / state / ids.

Patterned on tests/sixtracklib/cuda/track/test_track_job_track_until_c99.cpp
(CPU result is the oracle; 1e-14 there) and the reference's own element
fixtures.
"""
import os

import numpy as np
import pytest

import oracle_lib
from helpers import particles_from_arrays, particles_dict, assert_particles_equal
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.cobject import get_object
from sixtracklib_amd.particles import ALL_FIELDS, REAL_FIELDS
from sixtracklib_amd import lattices, beam_elements as be

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _job(eb, pb, fma=False, **kw):
    from sixtracklib_amd.trackjob import TrackJob
    cfg = "fma=1" if fma else "fma=0"
    return TrackJob(eb, pb, config_str=cfg, **kw)


def _copy(cb):
    return CBuffer.frombytes(cb.tobytes())


def _pdict(cb, idx=0):
    from sixtracklib_amd.particles import particles_in
    return particles_dict(particles_in(cb, idx))


def _gpu_until(eb, pb, turns, fma=False):
    job = _job(eb, pb, fma=fma)
    job.track_until(turns)
    job.collect()
    return _pdict(pb)


FMA_REL = 1e-13
SCALE = {"x": 1e-3, "y": 1e-3, "px": 1e-5, "py": 1e-5, "zeta": 1e-2, "delta": 1e-4,
         "psigma": 1e-4, "s": 1.0}


def _assert_fma_close(got, want, what="", rel=FMA_REL, nominal=True):
    """|a - b| <= rel * max(|b|, field scale); the scale is the larger of a
    nominal magnitude and the field's rms over the particle set, so fields
    crossing zero are judged relative to their typical size.  With
    nominal=False the scale is the rms alone."""
    for f in REAL_FIELDS:
        a, b = np.asarray(got[f]), np.asarray(want[f])
        scale = max(SCALE.get(f, 0.0) if nominal else 0.0,
                    float(np.sqrt(np.mean(b * b))), 1e-300)
        dev = np.abs(a - b) / np.maximum(np.abs(b), scale)
        worst = float(np.max(dev)) if dev.size else 0.0
        assert worst <= rel, "%s %s: worst relative deviation %.3g > %.3g" % (
            what, f, worst, rel)
    for f in ("particle_id", "at_element", "at_turn", "state"):
        assert np.array_equal(got[f], want[f]), "%s %s differs" % (what, f)


def test_device_present():
    from sixtracklib_amd import stcommon as st
    assert st.st_Hip_num_devices() >= 1


def test_drift_fixture_bit_exact_on_gpu():
    g = np.load(os.path.join(GOLDEN, "drift_golden.npz"))
    eb = CBuffer()
    for L in g["lengths"]:
        be.Drift(cbuffer=eb, length=L)
    pb, _ = particles_from_arrays({f: g["init_" + f] for f in ALL_FIELDS})
    got = _gpu_until(eb, pb, 1)
    assert_particles_equal(got, {f: g["final_" + f] for f in ALL_FIELDS}, what="drift")


def test_driftexact_fixture_bit_exact_on_gpu():
    g = np.load(os.path.join(GOLDEN, "driftexact_golden.npz"))
    eb = CBuffer()
    for L in g["lengths"]:
        be.DriftExact(cbuffer=eb, length=L)
    pb, _ = particles_from_arrays({f: g["init_" + f] for f in ALL_FIELDS})
    got = _gpu_until(eb, pb, 1)
    assert_particles_equal(got, {f: g["final_" + f] for f in ALL_FIELDS}, what="driftexact")


@pytest.mark.parametrize("turns,sigma", [(1, 1e-3), (20, 1e-3), (100, 1e-3), (100, 4e-4)])
def test_fodo_ring_exact_bitwise(turns, sigma):
    """100 turns of 1,000 particles through the 402-element FODO ring is the
    CPU-runnable reference configuration, compared in full.  Its 100
    sextupoles leave a dynamic aperture of about 0.4 mm at this beta, so at
    sigma = 1 mm most particles are lost within 5 turns (the comparison then
    checks loss element, turn and frozen state); sigma = 0.4 mm keeps about
    half of them for all 100 turns."""
    eb = lattices.fodo_ring()
    pb, _ = lattices.gaussian_beam(1000, seed=0, sigma_x=sigma, sigma_y=sigma,
                                   sigma_px=sigma / 100, sigma_py=sigma / 100)
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, turns) == 0
    got = _gpu_until(eb, pb, turns)
    assert_particles_equal(got, _pdict(ob), what="fodo %d turns" % turns)


def test_fodo_ring_fma_one_turn():
    eb = lattices.fodo_ring()
    pb, _ = lattices.gaussian_beam(1000, seed=0)
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 1) == 0
    got = _gpu_until(eb, pb, 1, fma=True)
    _assert_fma_close(got, _pdict(ob), "fodo fma")


@pytest.mark.parametrize("max_sigma", [5.0, 20.0])
def test_lhc_like_ring_fma_one_turn(max_sigma):
    """The 20,000-element ring, one turn, FMA build against the oracle.

    zeta starts at 0 on the polar grid, so after one turn it holds only the
    path-length terms rvv - (1 + (xp^2 + yp^2)/2) summed over the drifts;
    their rounding (1 + t/2 near 1) must follow the reference's for zeta to
    agree to 1e-13 of its own rms, and so must the Cavity's energy update
    for delta.  Every field is judged against its rms alone."""
    eb = lattices.lhc_like_ring(seed=0)
    sig = lattices.twiss_sigma(beta=236.0)
    pb, _ = lattices.polar_grid(32, 16, sigma_x=sig, sigma_y=sig,
                                max_sigma=max_sigma, delta=2.7e-4)
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 1, threads=8) == 0
    got = _gpu_until(eb, pb, 1, fma=True)
    want = _pdict(ob)
    assert np.all(want["state"] == 1)
    _assert_fma_close(got, want, "lhc_like_ring fma", nominal=False)


def test_lhc_like_ring_exact_build_bitwise():
    """The 20,000-element ring in the exact build: 1,024 particles of the
    0-20 sigma polar grid for 40 turns, every field bit-identical to the
    oracle; about 100 of them (the outer ones, beyond the dynamic
    aperture) are lost, at 7 different turns."""
    eb = lattices.lhc_like_ring(seed=0)
    sig = lattices.twiss_sigma(beta=236.0)
    pb, _ = lattices.polar_grid(16, 64, sigma_x=sig, sigma_y=sig, max_sigma=20.0, delta=2.7e-4)
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 40, threads=16) == 0
    got = _gpu_until(eb, pb, 40, fma=False)
    want = _pdict(ob)
    lost = want["state"] == 0
    assert 50 < np.count_nonzero(lost) < 200
    assert len(np.unique(want["at_turn"][lost])) >= 5
    assert_particles_equal(got, want, what="lhc_like_ring exact")


def test_fodo_ring_losses_match_oracle():
    """Large amplitudes: particles lost at the LimitRect and the global
    aperture; at_element / at_turn / state must match exactly."""
    eb = lattices.fodo_ring()
    pb, _ = lattices.gaussian_beam(2048, seed=3, sigma_x=0.03, sigma_y=0.02,
                                   sigma_px=3e-4, sigma_py=3e-4)
    ob = _copy(pb)
    oracle_lib.track_until(ob, eb, 10)
    got = _gpu_until(eb, pb, 10)
    want = _pdict(ob)
    assert np.count_nonzero(want["state"] == 0) > 100
    assert_particles_equal(got, want, what="fodo losses")


def test_fodo_ring_losses_in_steps_and_compaction():
    """track_until in steps (3, 7, 10 turns) gathers the still-active
    particles before each launch; the result equals one oracle call to 10
    turns bit for bit, with the gathering off, and with every request split
    into launches of 2 turns (list rebuilt between them)."""
    eb = lattices.fodo_ring()
    pb, _ = lattices.gaussian_beam(3000, seed=5, sigma_x=4e-4, sigma_y=4e-4,
                                   sigma_px=4e-6, sigma_py=4e-6)
    ob, pb2, pb3 = _copy(pb), _copy(pb), _copy(pb)
    oracle_lib.track_until(ob, eb, 10)
    from sixtracklib_amd.trackjob import TrackJob
    got = {}
    pb4 = _copy(pb)
    pb5 = _copy(pb)
    for key, buf, cfg in (("on", pb, "fma=0"), ("off", pb2, "fma=0 compact=0"),
                          ("chunked", pb3, "fma=0 turns_per_launch=2"),
                          ("pairs", pb4, "fma=0 ppl=2 turns_per_launch=3"),
                          ("lds", pb5, "fma=0 staging=lds turns_per_launch=3")):
        job = TrackJob(eb, buf, config_str=cfg)
        for t in (3, 7, 10):
            assert job.track_until(t).last_track_status == 0
        job.collect()
        got[key] = _pdict(buf)
    want = _pdict(ob)
    assert 100 < np.count_nonzero(want["state"] == 0) < 2900
    assert_particles_equal(got["on"], want, what="stepped, gathered")
    assert_particles_equal(got["off"], want, what="stepped, not gathered")
    assert_particles_equal(got["chunked"], want, what="stepped, 2 turns per launch")
    assert_particles_equal(got["pairs"], want, what="stepped, two particles per lane")
    assert_particles_equal(got["lds"], want, what="stepped, lattice staged through LDS")


@pytest.mark.parametrize("fma", [False, True])
def test_two_particles_per_lane_match_one(fma):
    """The two-particles-per-lane kernel gives every particle the result of
    the one-particle kernel bit for bit (exact and FMA builds), on a ring
    with losses, a particle count that leaves the last tile part-filled,
    and a Cavity, XYShift and SRotation in the lean lattice."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.lhc_like_ring(seed=3, n_cells=8, tune=8.31 * 8 / 40)
    sig = lattices.twiss_sigma(beta=236.0)
    pb, _ = lattices.polar_grid(45, 64, sigma_x=sig, sigma_y=sig, max_sigma=200.0,
                                delta=2.7e-4)
    pa = _copy(pb)
    ob = _copy(pb)
    res = {}
    for key, buf, cfg in (("one", pa, "ppl=1"), ("two", pb, "ppl=2")):
        job = TrackJob(eb, buf, config_str="fma=%d %s" % (int(fma), cfg))
        for t in (4, 9, 20):
            assert job.track_until(t).last_track_status == 0
        job.collect()
        res[key] = _pdict(buf)
    assert 50 < np.count_nonzero(res["one"]["state"] == 0) < 2800
    assert_particles_equal(res["two"], res["one"], what="pairs vs single")
    if not fma:
        oracle_lib.track_until(ob, eb, 20)
        assert_particles_equal(res["two"], _pdict(ob), what="pairs vs oracle")


def test_two_particles_per_lane_after_track_line_match_oracle():
    """Particles that start a track_until mid-lattice (s != 0 after a
    track_line without finishing the turn) run through the one-particle
    loop inside the two-particle kernel; every third particle also starts
    at a later turn.  Exact build against the one-particle kernel and the
    oracle, bit for bit."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.particles import particles_in
    eb = lattices.lhc_like_ring(seed=3, n_cells=8, tune=8.31 * 8 / 40)
    sig = lattices.twiss_sigma(beta=236.0)
    pb, _ = lattices.polar_grid(40, 64, sigma_x=sig, sigma_y=sig, max_sigma=80.0,
                                delta=2.7e-4)
    p = particles_in(pb)
    at = np.array(p.at_turn)
    at[::3] = 2
    p.at_turn[:] = at
    bufs = {"one": _copy(pb), "two": _copy(pb), "oracle": _copy(pb)}
    for key, cfg in (("one", "ppl=1"), ("two", "ppl=2")):
        job = TrackJob(eb, bufs[key], config_str="fma=0 " + cfg)
        job.track_line(0, 137, False)
        assert job.track_until(6).last_track_status == 0
        job.collect()
    oracle_lib.track_line(bufs["oracle"], eb, 0, 137, False)
    oracle_lib.track_until(bufs["oracle"], eb, 6)
    one, two = _pdict(bufs["one"]), _pdict(bufs["two"])
    assert 50 < np.count_nonzero(one["state"] == 0) < 2500
    assert_particles_equal(two, one, what="pairs vs single after track_line")
    assert_particles_equal(two, _pdict(bufs["oracle"]), what="pairs vs oracle after track_line")


@pytest.mark.parametrize("fma", [False, True])
def test_ppl_auto_matches_one_particle_per_lane(fma):
    """ppl=auto (the default) takes two particles per lane in the FMA build
    (the exact build keeps one) for a request whose particles fill at least
    70 % of the resident waves with
    128-particle tiles (22,938 with waves_per_cu=1, 256 resident waves).
    40,000 particles to 200 sigma start above it and fall below it in the
    first request (8,403 left after 3 turns); 35,840 to 20 sigma stay above
    it for the three requests.  Results equal the one-particle kernel's bit
    for bit."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.lhc_like_ring(seed=3, n_cells=8, tune=8.31 * 8 / 40)
    sig = lattices.twiss_sigma(beta=236.0)
    for n_amp, max_sigma in ((625, 200.0), (560, 20.0)):
        pb, _ = lattices.polar_grid(n_amp, 64, sigma_x=sig, sigma_y=sig, max_sigma=max_sigma,
                                    delta=2.7e-4)
        res = {}
        for key, cfg in (("auto", ""), ("one", " ppl=1")):
            buf = _copy(pb)
            job = TrackJob(eb, buf, config_str="fma=%d waves_per_cu=1%s" % (int(fma), cfg))
            for t in (3, 7, 12):
                assert job.track_until(t).last_track_status == 0
            job.collect()
            res[key] = _pdict(buf)
        if max_sigma > 100:
            assert np.count_nonzero(res["one"]["state"] == 0) > 30000
        assert_particles_equal(res["auto"], res["one"], what="auto vs ppl=1, %d particles" % (64 * n_amp))


@pytest.mark.parametrize("n_amp,max_sigma,later,fma", [
    (9375, 20.0, False, 1), (9375, 200.0, False, 1), (9375, 20.0, True, 1), (12500, 20.0, False, 1),
    (16000, 20.0, False, 1), (7500, 20.0, False, 1), (6250, 20.0, False, 1), (5700, 20.0, False, 1),
    (2350, 20.0, False, 1),
    (9375, 200.0, False, 0), (16000, 20.0, False, 0), (7500, 20.0, False, 0)])
def test_mixed_launches_match_one_particle_per_lane(n_amp, max_sigma, later, fma):
    """With ppl=auto (either build), a request of one particle set whose
    active particles sit at one turn runs mixed launches: some waves of
    every SIMD take two particles per lane, the others share the rest
    evenly, one per lane.  With 8,192 resident waves: from 524,289 to
    983,040 particles (9 to 15 lanes of 64 per SIMD) m pairs of w waves per
    SIMD, w = ceil((j + 1) / 2); up to 1,048,576 all 8 waves pairs; from
    65,537 to 524,288 (2 to 8 waves per SIMD one per lane) j waves per
    SIMD, or 6 with two pairs for 8.  On a 400-element ring: 600,000
    particles (4 pairs of 6 waves) to 20 sigma (all requests mixed), to 200
    sigma (most lost in the first turns, the later requests on smaller
    grids), and with every third particle starting two turns later (the
    first request not mixed, the next ones mixed); 800,000 particles (6
    pairs of 7 waves), 1,024,000 (every wave a pair wave, the particles
    spread evenly), and 480,000, 400,000, 364,800 and 150,400 particles (6
    waves with 2 pairs, 7, 6 and 3 waves per SIMD) to 20 sigma; in the
    exact build the 200-sigma, 1,024,000 and 480,000 cases.  Results equal
    the one-particle kernel's bit for bit."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.particles import particles_in
    eb = lattices.lhc_like_ring(seed=3, n_cells=8, tune=8.31 * 8 / 40)
    sig = lattices.twiss_sigma(beta=236.0)
    pb, _ = lattices.polar_grid(n_amp, 64, sigma_x=sig, sigma_y=sig, max_sigma=max_sigma,
                                delta=2.7e-4)
    if later:
        p = particles_in(pb)
        at = np.array(p.at_turn)
        at[::3] = 2
        p.at_turn[:] = at
    res = {}
    for key, cfg in (("mixed", "fma=%d" % fma), ("one", "fma=%d ppl=1" % fma)):
        buf = _copy(pb)
        job = TrackJob(eb, buf, config_str=cfg)
        for t in (3, 7, 12):
            assert job.track_until(t).last_track_status == 0
        job.collect()
        res[key] = _pdict(buf)
    lost = np.count_nonzero(res["one"]["state"] == 0)
    if max_sigma > 100:
        assert lost > 100000
    assert_particles_equal(res["mixed"], res["one"], what="mixed vs ppl=1")


@pytest.mark.parametrize("fma", [0, 1])
def test_mix_off_matches_mixed_launches(fma):
    """mix=0 turns the mixed launches off (ppl=auto then launches 64- or
    128-particle tiles, one wave each): 480,000 particles on a 400-element
    ring to 200 sigma, where the default takes mixed launches of every kind
    as particles are lost.  Results are identical bit for bit."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.lhc_like_ring(seed=3, n_cells=8, tune=8.31 * 8 / 40)
    sig = lattices.twiss_sigma(beta=236.0)
    pb, _ = lattices.polar_grid(7500, 64, sigma_x=sig, sigma_y=sig, max_sigma=200.0, delta=2.7e-4)
    res = {}
    for key, cfg in (("mixed", "fma=%d" % fma), ("off", "fma=%d mix=0" % fma)):
        buf = _copy(pb)
        job = TrackJob(eb, buf, config_str=cfg)
        for t in (3, 7, 12):
            assert job.track_until(t).last_track_status == 0
        job.collect()
        res[key] = _pdict(buf)
    assert np.count_nonzero(res["off"]["state"] == 0) > 50000
    assert_particles_equal(res["mixed"], res["off"], what="mixed vs mix=0")


@pytest.mark.parametrize("seed", [0, 1])
def test_kitchen_sink_exact_bitwise(seed):
    eb = lattices.kitchen_sink(seed)
    pb, _ = lattices.gaussian_beam(512, seed=seed, sigma_x=2e-3, sigma_y=2e-3,
                                   sigma_px=1e-4, sigma_py=1e-4, sigma_zeta=0.05)
    ob = _copy(pb)
    st_o = oracle_lib.track_until(ob, eb, 5)
    job = _job(eb, pb)
    st_g = job.track_until(5).last_track_status
    job.collect()
    assert st_o == st_g == 0
    assert_particles_equal(_pdict(pb), _pdict(ob), what="kitchen sink")


def test_kitchen_sink_fma_one_turn():
    eb = lattices.kitchen_sink(0)
    pb, _ = lattices.gaussian_beam(512, seed=5, sigma_x=2e-3, sigma_y=2e-3,
                                   sigma_px=1e-4, sigma_py=1e-4, sigma_zeta=0.05)
    ob = _copy(pb)
    oracle_lib.track_until(ob, eb, 1)
    got = _gpu_until(eb, pb, 1, fma=True)
    _assert_fma_close(got, _pdict(ob), "kitchen sink fma")


def test_track_line_lost_particle_behaviour():
    """tests/sixtracklib/common/track/test_track_job_track_line_c99.cpp:
    237-384: DriftExact 5 m, quad knl[1]=1e-2 L=1, LimitRect +-0.1,
    DriftExact 5 m; 100 particles x = 0.002 i; 2 turns tracked by lines
    equal the until-turn result."""
    def lattice():
        eb = CBuffer()
        be.DriftExact(cbuffer=eb, length=5.0)
        be.Multipole(cbuffer=eb, knl=[0.0, 1e-2], length=1.0)
        be.LimitRect(cbuffer=eb, min_x=-0.1, max_x=0.1, min_y=-0.1, max_y=0.1)
        be.DriftExact(cbuffer=eb, length=5.0)
        return eb
    x = 0.002 * np.arange(100)
    pb_line, _ = lattices.make_particles(100, p0c=1e9, x=x, px=1e-3 * np.arange(100) / 100)
    pb_until = _copy(pb_line)
    pb_oracle = _copy(pb_line)
    eb1, eb2 = lattice(), lattice()
    job = _job(eb1, pb_line)
    for _ in range(2):
        job.track_line(0, 2)
        job.track_line(2, 3)
        job.track_line(3, 4, finish_turn=True)
    job.collect()
    job2 = _job(eb2, pb_until)
    job2.track_until(2)
    job2.collect()
    assert_particles_equal(_pdict(pb_line), _pdict(pb_until), what="line vs until")
    oracle_lib.track_until(pb_oracle, lattice(), 2)
    assert_particles_equal(_pdict(pb_until), _pdict(pb_oracle), what="until vs oracle")
    assert np.count_nonzero(_pdict(pb_until)["state"] == 0) > 0


def test_track_line_matches_oracle_segments():
    eb = lattices.kitchen_sink(1)
    n = eb.n_objects
    pb, _ = lattices.gaussian_beam(256, seed=9, sigma_x=2e-3, sigma_y=2e-3)
    ob = _copy(pb)
    job = _job(eb, pb)
    cuts = [0, 3, 7, 12, n]
    for a, b in zip(cuts[:-1], cuts[1:]):
        job.track_line(a, b, finish_turn=(b == n))
        oracle_lib.track_line(ob, eb, a, b, b == n)
    job.collect()
    assert_particles_equal(_pdict(pb), _pdict(ob), what="line segments")


def test_elem_by_elem_output_matches_oracle():
    eb = lattices.fodo_ring(n_cells=10)
    pb, _ = lattices.gaussian_beam(200, seed=4, sigma_x=0.02, sigma_y=0.02)
    ob = _copy(pb)
    from sixtracklib_amd.trackjob import TrackJob
    job = TrackJob(eb, pb, until_turn_elem_by_elem=3, config_str="fma=0")
    assert job.has_elem_by_elem_output
    job.track_elem_by_elem(3)
    job.collect()
    out = job.output
    # oracle with its own output set and the same config
    from sixtracklib_amd.particles import particles_in, Particles
    n_out = particles_in(out, job.elem_by_elem_output_offset).num_particles
    ocb = CBuffer(max_slots=22 * n_out + 64)
    Particles(cbuffer=ocb, num_particles=n_out)
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=200,
                               num_elements_to_store=eb.n_objects, num_turns_to_store=3,
                               min_particle_id=0, min_element_id=0, min_turn=0,
                               max_particle_id=199, max_element_id=eb.n_objects - 1,
                               max_turn=2, is_rolling=1,
                               out_store_addr=ocb.get_object_address(0))
    oracle_lib.track_elem_by_elem(ob, eb, cfg, 3)
    assert_particles_equal(_pdict(pb), _pdict(ob), what="ebe particles")
    assert_particles_equal(_pdict(out, job.elem_by_elem_output_offset), _pdict(ocb, 0),
                           what="ebe store")


def _collective_beam(n=192):
    """Matched beam for the short collective ring; the first 12 particles
    sit at growing amplitudes so that some are lost in turns 0 and 1
    (n >= 12)."""
    _, p = lattices.matched_gaussian_beam(n, seed=3)
    x = np.array(p.x)
    x[:12] = np.linspace(0.004, 0.0098, 12)
    pb, _ = lattices.make_particles(n, x=x, px=np.array(p.px), y=np.array(p.y),
                                    py=np.array(p.py), zeta=np.array(p.zeta),
                                    delta=np.array(p.delta))
    return pb


def test_elem_by_elem_reference_store_rolls_over_turns():
    """The reference's own store (particles x elements x 2 turns, rolling)
    for 5 turns of element-by-element tracking: turns 2-4 overwrite the
    slots of turns 0-1, exactly as the oracle does with the same
    configuration."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.particles import particles_in, Particles
    eb = lattices.collective_ring(n_cells=8, n_sc=4)
    E = eb.n_objects
    pb = _collective_beam(130)
    P = particles_in(pb).num_particles
    ob = _copy(pb)
    job = TrackJob(eb, pb, until_turn_elem_by_elem=2, config_str="fma=0")
    off = job.elem_by_elem_output_offset
    job.track_elem_by_elem(5)
    job.collect()
    got = _pdict(job.output, off)
    assert len(got["x"]) == P * E * 2
    ocb = CBuffer(max_slots=22 * P * E * 2 + 64)
    Particles(cbuffer=ocb, num_particles=P * E * 2)
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=P, num_elements_to_store=E,
                               num_turns_to_store=2, min_particle_id=0, min_element_id=0,
                               min_turn=0, max_particle_id=P - 1, max_element_id=E - 1,
                               max_turn=1, is_rolling=1, out_store_addr=ocb.get_object_address(0))
    assert oracle_lib.track_elem_by_elem(ob, eb, cfg, 5) == 0
    assert_particles_equal(_pdict(pb), _pdict(ob), what="particles")
    assert_particles_equal(got, _pdict(ocb), what="rolling reference-size store")
    assert np.any(got["at_turn"] >= 3)


@pytest.mark.parametrize("n,turns,shuffle", [(130, 2, False), (129, 2, False),
                                             (131, 3, False), (130, 2, True)])
def test_elem_by_elem_store_layouts_match_oracle(n, turns, shuffle):
    """The reference-size element-by-element store for particle counts and
    ids that change how a wave's records lie in the store, against the
    oracle's store bit for bit: 130 particles (neighbouring particles 8 and
    9 lost at different elements), 129 (odd: consecutive passes of a wave
    alternate between even and odd record offsets; the last wave holds one
    particle), 131 over 3 turns (odd capacity: the field arrays alternate
    between 16-byte and 8-byte alignment), and 130 with shuffled particle
    ids (neighbouring lanes' records are not adjacent)."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.particles import particles_in, Particles
    eb = lattices.collective_ring(n_cells=8, n_sc=4)
    E = eb.n_objects
    pb = _collective_beam(n)
    if shuffle:
        pv = particles_in(pb)
        type(pv).__dict__["particle_id"]._view(pv)[:] = np.random.default_rng(5).permutation(n)
    ob = _copy(pb)
    job = TrackJob(eb, pb, until_turn_elem_by_elem=turns, config_str="fma=0")
    off = job.elem_by_elem_output_offset
    job.track_elem_by_elem(turns)
    job.collect()
    got = _pdict(job.output, off)
    assert len(got["x"]) == n * E * turns
    ocb = CBuffer(max_slots=22 * n * E * turns + 64)
    Particles(cbuffer=ocb, num_particles=n * E * turns)
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=n, num_elements_to_store=E,
                               num_turns_to_store=turns, min_particle_id=0, min_element_id=0,
                               min_turn=0, max_particle_id=n - 1, max_element_id=E - 1,
                               max_turn=turns - 1, is_rolling=1,
                               out_store_addr=ocb.get_object_address(0))
    assert oracle_lib.track_elem_by_elem(ob, eb, cfg, turns) == 0
    assert_particles_equal(_pdict(pb), _pdict(ob), what="particles")
    d = _pdict(ob)                                           # the pair (8, 9) breaks:
    assert d["state"][9] == 0 and (d["state"][8] == 1 or     # 9 is lost before 8
           (d["at_turn"][8], d["at_element"][8]) > (d["at_turn"][9], d["at_element"][9]))
    assert_particles_equal(got, _pdict(ocb), what="store")


def test_elem_by_elem_in_windows_matches_oracle():
    """A set larger than one wave per SIMD (70,000 particles; 65,536 fill
    the 1,024 SIMDs of a 256-CU device) is tracked element by element in
    one-turn launches over windows of 65,536 particles that rotate through
    the set (4 launches for 3 turns, windows wrapping past the end);
    records and particles equal the oracle's bit for bit, losses at the
    aperture included."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.particles import particles_in, Particles
    eb = CBuffer()
    be.Drift(cbuffer=eb, length=0.7)
    be.Multipole(cbuffer=eb, knl=[0.0, 0.02, 0.3], ksl=[0.0, 0.0, 0.05])
    be.LimitRect(cbuffer=eb, min_x=-0.01, max_x=0.01, min_y=-0.01, max_y=0.01)
    be.Drift(cbuffer=eb, length=1.3)
    E, T, n = 4, 3, 70000
    pb, _ = lattices.gaussian_beam(n, seed=8, sigma_x=4e-3, sigma_y=4e-3, sigma_px=2e-4,
                                   sigma_py=2e-4)
    ob, cb = _copy(pb), _copy(pb)
    job = TrackJob(eb, pb, until_turn_elem_by_elem=T, config_str="fma=0")
    off = job.elem_by_elem_output_offset
    assert job.track_elem_by_elem(T).last_track_status == 0
    job.collect()
    got = _pdict(job.output, off)
    # compact=0: no select pass, the live particles in rounds of one wave
    # per SIMD for the whole request
    job2 = TrackJob(eb, cb, until_turn_elem_by_elem=T, config_str="fma=0 compact=0")
    assert job2.track_elem_by_elem(T).last_track_status == 0
    job2.collect()
    got2 = _pdict(job2.output, job2.elem_by_elem_output_offset)
    ocb = CBuffer(max_slots=22 * n * E * T + 64)
    Particles(cbuffer=ocb, num_particles=n * E * T)
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=n, num_elements_to_store=E,
                               num_turns_to_store=T, min_particle_id=0, min_element_id=0,
                               min_turn=0, max_particle_id=n - 1, max_element_id=E - 1,
                               max_turn=T - 1, is_rolling=1,
                               out_store_addr=ocb.get_object_address(0))
    assert oracle_lib.track_elem_by_elem(ob, eb, cfg, T) == 0
    want = _pdict(ob)
    assert 1000 < np.count_nonzero(want["state"] == 0) < n - 1000
    assert_particles_equal(_pdict(pb), want, what="particles")
    assert_particles_equal(got, _pdict(ocb), what="store")
    assert_particles_equal(_pdict(cb), want, what="particles, compact=0")
    assert_particles_equal(got2, _pdict(ocb), what="store, compact=0")


@pytest.mark.parametrize("passes", [300, 407 * 4 - 5])
def test_elem_by_elem_rolling_passes_keep_the_latest_records(passes):
    """Collective ring (BB6D, BB4D, space charge; 407 elements), 4 turns
    element by element into a store of ``passes`` element passes per
    particle (ebe_store_passes): every slot holds the latest record of the
    oracle's full store that maps to it under the rolling index
    ((turn * E + element) mod K) * P + particle; unwritten slots keep their
    initial contents; the tracked particles are bit-identical."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.particles import particles_in, Particles
    eb = lattices.collective_ring(n_cells=8, n_sc=4)
    E, T = eb.n_objects, 4
    pb = _collective_beam()
    P = particles_in(pb).num_particles
    ob = _copy(pb)
    job = TrackJob(eb, pb, until_turn_elem_by_elem=T,
                   config_str="fma=0 ebe_store_passes=%d" % passes)
    off = job.elem_by_elem_output_offset
    init = _pdict(job.output, off)
    assert len(init["x"]) == P * passes
    job.track_elem_by_elem(T)
    job.collect()
    got = _pdict(job.output, off)
    # oracle: the reference's full store, P x E x T records
    n_full = P * E * T
    fcb = CBuffer(max_slots=22 * n_full + 64)
    Particles(cbuffer=fcb, num_particles=n_full)
    fv = particles_in(fcb)
    type(fv).__dict__["particle_id"]._view(fv)[:] = -1
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=P, num_elements_to_store=E,
                               num_turns_to_store=T, min_particle_id=0, min_element_id=0,
                               min_turn=0, max_particle_id=P - 1, max_element_id=E - 1,
                               max_turn=T - 1, is_rolling=1,
                               out_store_addr=fcb.get_object_address(0))
    assert oracle_lib.track_elem_by_elem(ob, eb, cfg, T) == 0
    assert_particles_equal(_pdict(pb), _pdict(ob), what="ebe particles")
    full = particles_dict(fv)
    want = {f: np.array(v) for f, v in init.items()}
    lane = np.arange(P)
    for j in range(E * T):
        src = slice(j * P, (j + 1) * P)
        written = full["particle_id"][src] >= 0
        dst = (j % passes) * P + lane[written]
        for f in want:
            want[f][dst] = full[f][src][written]
    assert np.count_nonzero(full["particle_id"] >= 0) > P * E * (T - 1)
    assert_particles_equal(got, want, what="rolling store")


@pytest.mark.parametrize("method", ["linear", "cubic"])
def test_sc_interpolated_tables_from_stored_buffer(method):
    """Line-density tables the reference's way
    (examples/python/spacecharge/line_density_profile.py): tables in a
    stored buffer, address assignments to two SCInterpolatedProfile
    elements, assign_all_addresses; the GPU result is bit-identical to the
    oracle tracking the same (now pointing) host lattice, and differs from
    the fallback density.  A changed table reaches the device with
    push_stored_buffer."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.sc_line_density import (LineDensityProfileData,
                                                 LineDensityProfileData_create_buffer_assignment)
    eb = lattices.fodo_ring(n_cells=4)
    sc_idx = []
    for sx, sy in ((1e-3, 1.3e-3), (1.1e-3, 1.1e-3)):
        sc_idx.append(eb.n_objects)
        be.SCInterpolatedProfile(cbuffer=eb, number_of_particles=1e13, sigma_x=sx, sigma_y=sy,
                                 length=50.0, line_density_prof_fallback=0.5)
        be.Drift(cbuffer=eb, length=0.5)
    tb = CBuffer()
    z = np.linspace(-0.5, 0.5, 101)
    for sz in (0.08, 0.12):
        prof = LineDensityProfileData(cbuffer=tb, method=method, capacity=len(z), z0=-0.5,
                                      dz=z[1] - z[0], values=np.exp(-z ** 2 / (2 * sz * sz)))
        prof.prepare_interpolation()
    pb, _ = lattices.gaussian_beam(512, seed=5, sigma_zeta=0.1)
    ob, fb = _copy(pb), _copy(pb)
    oracle_lib.track_until(fb, eb, 3)                     # fallback densities
    job = _job(eb, pb)
    bid = job.add_stored_buffer(cbuffer=tb)
    assert bid == 1024 and job.num_stored_buffers == 1
    for k, i in enumerate(sc_idx):
        assert LineDensityProfileData_create_buffer_assignment(job, i, bid, k)
    assert not LineDensityProfileData_create_buffer_assignment(job, sc_idx[0], bid, 0)
    job.commit_address_assignments()
    job.assign_all_addresses()
    for k, i in enumerate(sc_idx):
        assert get_object(eb, i).interpol_data_addr == tb.get_object_address(k)
    job.track_until(3)
    job.collect()
    oracle_lib.track_until(ob, eb, 3)
    assert_particles_equal(_pdict(pb), _pdict(ob), what="tables")
    assert not np.array_equal(_pdict(pb)["px"], _pdict(fb)["px"])
    # new table values: pushed, then tracked on from turn 3
    from sixtracklib_amd.particles import particles_in
    prof0 = get_object(tb, 0)
    prof0.values = np.array(prof0.values) * 3.0
    prof0.prepare_interpolation()
    job.push_stored_buffer(bid)
    job.track_until(5)
    job.collect()
    oracle_lib.track_until(ob, eb, 5)
    assert_particles_equal(_pdict(pb), _pdict(ob), what="pushed tables")
    assert particles_in(pb).num_particles == 512


def _monitor_case(monitors, n_particles, turns):
    """Track with BeamMonitors on the GPU and in the oracle (each with its
    own output sets, same particle range) -> (gpu particles, gpu outputs,
    oracle particles, oracle outputs, gpu status, oracle status)."""
    from sixtracklib_amd.trackjob import TrackJob
    from sixtracklib_amd.particles import Particles
    from sixtracklib_amd.cobject import get_object
    from sixtracklib_amd import stcommon as st

    def lattice():
        eb = lattices.fodo_ring(n_cells=20)
        for kw in monitors:
            be.BeamMonitor(cbuffer=eb, **kw)
        return eb
    eb = lattice()
    pb, _ = lattices.gaussian_beam(n_particles, seed=11, sigma_x=0.02, sigma_y=0.02)
    ob = _copy(pb)
    job = TrackJob(eb, pb, config_str="fma=0")
    assert job.has_beam_monitor_output and job.num_beam_monitors == len(monitors)
    st_g = st.st_TrackJob_track_until(job.ptr_st_track_job, turns)
    job.collect()
    out, off = job.output, job.beam_monitor_output_offset
    oeb = lattice()
    m_idx = [i for i in range(oeb.n_objects) if oeb.get_object_typeid(i) == 10]
    ocb = CBuffer(max_slots=22 * 64 + 21 * n_particles * sum(
        int(get_object(oeb, i).num_stores) for i in m_idx) + 64)
    for i in m_idx:
        Particles(cbuffer=ocb, num_particles=int(get_object(oeb, i).num_stores) * n_particles)
    for k, i in enumerate(m_idx):
        m = get_object(oeb, i)
        m.min_particle_id = 0
        m.max_particle_id = n_particles - 1
        m.out_address = ocb.get_object_address(k)
    st_o = oracle_lib.track_until(ob, oeb, turns)
    g_out = [_pdict(out, off + k) for k in range(len(m_idx))]
    o_out = [_pdict(ocb, k) for k in range(len(m_idx))]
    return _pdict(pb), g_out, _pdict(ob), o_out, st_g, st_o


def test_beam_monitor_turn_ordered_matches_oracle():
    mons = [dict(num_stores=5, start=1, skip=2, is_rolling=1, is_turn_ordered=1),
            dict(num_stores=3, start=0, skip=1, is_rolling=0, is_turn_ordered=1)]
    gp, go, op, oo, st_g, st_o = _monitor_case(mons, 300, 12)
    assert st_g == st_o == 0
    assert_particles_equal(gp, op, what="monitor particles")
    for k in range(2):
        assert_particles_equal(go[k], oo[k], what="monitor %d store" % k)


def test_beam_monitor_particle_ordered_matches_oracle():
    """Non-turn-ordered stores go to pid_off * num_particles + store_idx
    (be_monitor/track.h:105-106); with num_stores == num_particles every
    slot is written by exactly one (particle, turn)."""
    mons = [dict(num_stores=8, start=0, skip=1, is_rolling=0, is_turn_ordered=0)]
    gp, go, op, oo, st_g, st_o = _monitor_case(mons, 8, 10)
    assert st_g == st_o == 0
    assert_particles_equal(gp, op, what="particles")
    assert_particles_equal(go[0], oo[0], what="particle-ordered store")


def test_beam_monitor_out_of_range_store_marks_lost_like_reference():
    """With fewer stores than particles the particle-ordered index runs
    past the output set: the reference's copy fails, the particle is
    marked lost and the status is -1 (be_monitor/track.h:160-166)."""
    mons = [dict(num_stores=3, start=0, skip=1, is_rolling=0, is_turn_ordered=0)]
    gp, go, op, oo, st_g, st_o = _monitor_case(mons, 300, 4)
    assert st_g == st_o == -1
    assert np.count_nonzero(op["state"] == 0) > 250
    assert_particles_equal(gp, op, what="particles after failed stores")
    assert_particles_equal(go[0], oo[0], what="store")


def test_bbsimple_element_by_element_on_gpu():
    """Reference fixture dump tracked element by element with track_line."""
    from test_oracle_golden import _bbsimple_lattice
    g = np.load(os.path.join(GOLDEN, "bbsimple_golden.npz"))
    eb = _bbsimple_lattice(g)
    for ii in range(eb.n_objects):
        start = {f: g["p%d_%s" % (ii, f)] for f in ALL_FIELDS}
        want = {f: g["p%d_%s" % (ii + 1, f)] for f in ALL_FIELDS}
        pb, _ = particles_from_arrays(start)
        job = _job(eb, pb)
        job.track_line(ii, ii + 1)
        job.collect()
        assert_particles_equal(_pdict(pb), want, abs_tol=1e-15, check_index=False,
                               what="bbsimple element %d" % ii)

"""Seeded random cases, exact build against the oracle, bit for bit.

Each case draws a lattice (a lean one of Drift, DriftExact, Multipole of
orders 0-14 with and without curvature, Cavity, XYShift, SRotation,
DipoleEdge and the three apertures, or with monitors, beam-beam, space
charge and RF multipoles added), a ragged particle set (1 to 3,000
particles, some lost on entry, some starting at later turns), a job
configuration (particles per lane, lattice staging, balanced or chunked
scheduling, gathering of live particles, turns per launch, block size,
queues, wave priority, resident waves) and a split of the request into
track_until steps.  Apertures are sized so that part of each set is
lost.  The GPU result equals the oracle's in every field.  The case number
names the seed.
"""
import math

import numpy as np
import pytest

import oracle_lib
from helpers import particles_dict, assert_particles_equal
from sixtracklib_amd import beam_elements as be
from sixtracklib_amd import lattices
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import particles_in

pytestmark = pytest.mark.gpu

N_CASES = 64
N_EBE_CASES = 16


def _lattice(rng, heavy, max_elements=500):
    n = int(rng.integers(8, max_elements))
    eb = CBuffer(max_slots=n * 64 + 4096, max_objects=n + 64, max_pointers=n + 64)
    fact = [math.factorial(k) for k in range(15)]
    for _ in range(n):
        r = rng.random()
        if r < 0.40:
            be.Drift(cbuffer=eb, length=float(rng.uniform(0.0, 0.5)))
        elif r < 0.75:
            order = int(rng.integers(0, 15)) if rng.random() < 0.3 else int(rng.integers(0, 4))
            scale = np.array([3e-7 * fact[k] / 0.005 ** k for k in range(order + 1)])
            kw = {}
            if rng.random() < 0.15:
                kw = dict(hxl=float(rng.normal(0, 1e-5)), hyl=float(rng.normal(0, 5e-6)),
                          length=float(rng.uniform(0.5, 3.0)))
            be.Multipole(cbuffer=eb, knl=rng.normal(0, 1, order + 1) * scale,
                         ksl=rng.normal(0, 1, order + 1) * scale * (rng.random() < 0.5), **kw)
        elif r < 0.79:
            be.DriftExact(cbuffer=eb, length=float(rng.uniform(0.0, 0.5)))
        elif r < 0.83:
            be.XYShift(cbuffer=eb, dx=float(rng.normal(0, 1e-5)), dy=float(rng.normal(0, 1e-5)))
        elif r < 0.87:
            be.SRotation(cbuffer=eb, angle=float(rng.uniform(-180, 180)))
        elif r < 0.89:
            be.DipoleEdge(cbuffer=eb, r21=float(rng.normal(0, 1e-4)), r43=float(rng.normal(0, 1e-4)))
        elif r < 0.91:
            be.Cavity(cbuffer=eb, voltage=float(rng.uniform(0, 1e7)),
                      frequency=float(rng.uniform(1e8, 8e8)), lag=float(rng.uniform(0, 360)))
        elif r < 0.94:
            a = float(rng.uniform(3e-3, 1e-2))
            be.LimitRect(cbuffer=eb, min_x=-a, max_x=a * 1.2, min_y=-a * 0.8, max_y=a)
        elif r < 0.97:
            be.LimitEllipse(cbuffer=eb, a=float(rng.uniform(3e-3, 1e-2)),
                            b=float(rng.uniform(3e-3, 1e-2)))
        else:
            be.LimitRectEllipse(cbuffer=eb, max_x=6e-3, max_y=5e-3, a_squ=3e-5, b_squ=2e-5)
        if heavy and rng.random() < 0.02:
            k = int(rng.integers(0, 5))
            if k == 0:
                be.BeamMonitor(cbuffer=eb, num_stores=int(rng.integers(1, 5)),
                               start=int(rng.integers(0, 3)), skip=int(rng.integers(1, 3)),
                               is_rolling=int(rng.integers(0, 2)))
            elif k == 1:
                be.BeamBeam4D(cbuffer=eb, charge=1e11, sigma_x=2e-3,
                              sigma_y=float(rng.uniform(1e-3, 2e-3)), beta_r=1.0,
                              min_sigma_diff=1e-10, x_bb=1e-4, y_bb=-1e-4, d_px=0.0, d_py=0.0,
                              enabled=1.0)
            elif k == 2:
                be.SCQGaussProfile(cbuffer=eb, number_of_particles=1e11, bunchlength_rms=0.1,
                                   sigma_x=1.3e-3, sigma_y=float(rng.uniform(0.8e-3, 1.3e-3)),
                                   length=1.0, q_param=float(rng.choice([0.8, 0.9, 1.0, 1.2, 1.4])))
            elif k == 3:
                be.SCCoasting(cbuffer=eb, number_of_particles=1e11, circumference=1000.0,
                              sigma_x=1.2e-3, sigma_y=0.9e-3, length=1.0, x_co=0.0, y_co=0.0)
            else:
                be.RFMultipole(cbuffer=eb, voltage=1e5, frequency=400e6, lag=10.0,
                               knl=[1e-6, 1e-3], ksl=[0.0, 2e-4], pn=[0.0, 30.0], ps=[10.0, 0.0])
    return eb


def _particles(rng, max_particles=3000, later_starts=True):
    n = int(rng.integers(1, max_particles + 1))
    s = float(rng.uniform(5e-4, 2.5e-3))
    pb, _ = lattices.gaussian_beam(n, seed=int(rng.integers(0, 1 << 30)), sigma_x=s, sigma_y=s,
                                   sigma_px=s / 1000, sigma_py=s / 1000, sigma_zeta=0.05,
                                   sigma_delta=2e-4)
    p = particles_in(pb)
    state = np.array(p.state)
    at_turn = np.array(p.at_turn)
    lost = rng.random(n) < 0.05
    state[lost] = 0
    later = (rng.random(n) < 0.1) & later_starts
    at_turn[later] = rng.integers(1, 3, int(np.count_nonzero(later)))
    p.state[:] = state
    p.at_turn[:] = at_turn
    return pb


def _config(rng, heavy):
    cfg = ["fma=0"]
    cfg.append("ppl=%d" % (2 if (not heavy and rng.random() < 0.25) else 1))
    if not heavy and rng.random() < 0.25:
        cfg.append("staging=lds")
    cfg.append("balance=%d" % int(rng.random() < 0.6))
    cfg.append("compact=%d" % int(rng.random() < 0.85))
    cfg.append("turns_per_launch=%d" % int(rng.choice([0, 0, 1, 3])))
    cfg.append("block=%d" % int(rng.choice([64, 128, 256])))
    cfg.append("queues=%d" % int(rng.choice([1, 8])))
    cfg.append("wave_prio=%d" % int(rng.random() < 0.7))
    if rng.random() < 0.3:
        cfg.append("waves_per_cu=%d" % int(rng.integers(1, 5)))
    return " ".join(cfg)


@pytest.mark.parametrize("case", range(N_CASES))
def test_random_case_matches_oracle(case):
    from sixtracklib_amd.trackjob import TrackJob
    rng = np.random.default_rng(1000 + case)
    heavy = bool(case % 4 == 3)
    eb = _lattice(rng, heavy)
    pb = _particles(rng)
    turns = int(rng.integers(1, 13))
    cfg = _config(rng, heavy)
    steps = sorted(set(int(t) for t in rng.integers(1, turns + 1, int(rng.integers(0, 3))))) + [turns]
    ob = CBuffer.frombytes(pb.tobytes())
    at0 = np.array(particles_in(ob).at_turn).copy()
    assert oracle_lib.track_until(ob, eb, turns, threads=16) == 0
    job = TrackJob(eb, pb, config_str=cfg)
    for t in steps:
        job.track_until(t)
    job.collect()
    got, want = particles_dict(particles_in(pb)), particles_dict(particles_in(ob))
    assert_particles_equal(got, want, what="case %d (%s; %d turns in steps %s)"
                           % (case, cfg, turns, steps))
    assert job.last_track_status == 0
    # the case tracked something and lost something, at more than one turn
    # when it ran several
    lost = want["state"] == 0
    assert np.sum(want["at_turn"] - at0) > 0
    assert np.any(lost) or len(lost) < 20
    if turns > 2:
        assert len(np.unique(want["at_turn"][lost])) >= 2


@pytest.mark.parametrize("case", range(N_EBE_CASES))
def test_random_elem_by_elem_case_matches_oracle(case):
    """Element by element on a random lattice into the reference-size store:
    the oracle runs with the job's own NS(ElemByElemConfig) (its index
    ranges and rolling flag) on a store of the same capacity; every record
    and every final particle are bit-identical."""
    import ctypes as ct
    from sixtracklib_amd import stcommon as st
    from sixtracklib_amd.particles import Particles
    from sixtracklib_amd.trackjob import TrackJob
    rng = np.random.default_rng(5000 + case)
    heavy = bool(case % 4 == 3)
    eb = _lattice(rng, heavy, max_elements=160)
    pb = _particles(rng, max_particles=300, later_starts=False)
    turns = int(rng.integers(1, 4))
    cfg = "fma=0 block=%d" % int(rng.choice([64, 128, 256]))
    ob = CBuffer.frombytes(pb.tobytes())
    job = TrackJob(eb, pb, until_turn_elem_by_elem=turns, config_str=cfg)
    jc = st.st_TrackJob_get_elem_by_elem_config(job.ptr_st_track_job).contents
    off = job.elem_by_elem_output_offset
    cap = particles_in(job.output, off).num_particles
    ocb = CBuffer(max_slots=22 * cap + 64)
    Particles(cbuffer=ocb, num_particles=cap)
    oc = oracle_lib.EbeConfig(**{f: getattr(jc, f) for f, _ in oracle_lib.EbeConfig._fields_
                                 if f != "out_store_addr"})
    oc.out_store_addr = ocb.get_object_address(0)
    assert oracle_lib.track_elem_by_elem(ob, eb, oc, turns) == 0
    job.track_elem_by_elem(turns)
    job.collect()
    what = "ebe case %d (%s; %d turns)" % (case, cfg, turns)
    assert_particles_equal(particles_dict(particles_in(job.output, off)),
                           particles_dict(particles_in(ocb)), what=what + " store")
    assert_particles_equal(particles_dict(particles_in(pb)), particles_dict(particles_in(ob)),
                           what=what + " particles")

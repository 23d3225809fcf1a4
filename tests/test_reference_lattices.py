"""The reference's LHC lattices (lhc_no_bb: 18,657 elements, multipoles of
orders 0-14; beambeam: 10,866 elements with 72 BeamBeam4D and 2
BeamBeam6D) rebuilt from the element tables of tests/golden/, tracked by
the oracle (CPU, no GPU needed).

The golden results were computed by the oracle's libm build on the
lattices read from the reference's own files; the rebuilt lattices must
give them bit for bit.  The stmath build (the GPU's transcendentals)
gives the same bits on lhc_no_bb; on beambeam the exponentials of the
72 beam-beam kicks round differently from glibc's in some calls, which
moves the coordinates by about 1e-11 of their rms after 2 turns.
"""
import numpy as np
import pytest

import oracle_lib
import reference_lattices as rl
from helpers import particles_dict, assert_particles_equal
from sixtracklib_amd.particles import particles_in, ALL_FIELDS, REAL_FIELDS


def _golden(t):
    return {f: t["final_" + f] for f in ALL_FIELDS}


@pytest.mark.parametrize("name,n_elem", [("lhc_no_bb", 18657), ("beambeam", 10866)])
def test_rebuilt_lattice_reproduces_golden_with_libm_oracle(name, n_elem):
    t = rl.load(name)
    eb = rl.rebuild(t)
    assert eb.n_objects == n_elem
    pb = rl.beam(64)
    init = particles_dict(particles_in(pb))
    assert_particles_equal(init, {f: t["init_" + f] for f in ALL_FIELDS}, what="initial")
    assert oracle_lib.track_until(pb, eb, int(t["turns"]), libm=True) == 0
    assert_particles_equal(particles_dict(particles_in(pb)), _golden(t), what=name)


def test_lhc_no_bb_multipole_orders_up_to_14():
    t = rl.load("lhc_no_bb")
    orders = t["t4_order"]
    assert orders.max() == 14
    for k in (11, 13, 14):
        assert np.count_nonzero(orders == k) > 100


def test_stmath_oracle_on_reference_lattices():
    t = rl.load("lhc_no_bb")
    pb = rl.beam(64)
    assert oracle_lib.track_until(pb, rl.rebuild(t), 2, libm=False) == 0
    assert_particles_equal(particles_dict(particles_in(pb)), _golden(t), what="lhc_no_bb stmath")
    t = rl.load("beambeam")
    pb = rl.beam(64)
    assert oracle_lib.track_until(pb, rl.rebuild(t), 2, libm=False) == 0
    got, want = particles_dict(particles_in(pb)), _golden(t)
    for f in ("particle_id", "at_element", "at_turn", "state"):
        assert np.array_equal(got[f], want[f])
    for f in REAL_FIELDS:
        b = want[f]
        scale = max(float(np.sqrt(np.mean(b * b))), 1e-300)
        assert np.max(np.abs(got[f] - b)) <= 1e-10 * scale, f

"""Pin the CPU oracle against the reference's own golden data.

Fixtures (tests/golden/*.npz) were extracted from the reference's
tests/testdata by tests/golden/make_fixtures.py:
* Drift / DriftExact: tests/sixtracklib/common/test_track_c99.cpp:298-357
  compares at machine epsilon; the oracle is required to be bit-exact.
* bbsimple element-by-element: test_track_c99.cpp:647-665 at 1e-15 abs.
* cerrf on the Faddeeva grid: reference cerrf agrees with the grid to
  4.5e-10 relative.
The libm build is the reference's arithmetic; the stmath build (the one
the GPU is compared against bit-for-bit) must meet the same tolerances.
"""
import os

import numpy as np
import pytest

import oracle_lib
from helpers import particles_from_arrays, particles_dict, assert_particles_equal
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd import beam_elements as be
from sixtracklib_amd.particles import ALL_FIELDS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.mark.parametrize("libm", [True, False])
@pytest.mark.parametrize("name,cls", [("drift_golden.npz", be.Drift),
                                      ("driftexact_golden.npz", be.DriftExact)])
def test_drift_fixture_bit_exact(name, cls, libm):
    g = _load(name)
    eb = CBuffer()
    for L in g["lengths"]:
        cls(cbuffer=eb, length=L)
    pb, p = particles_from_arrays({f: g["init_" + f] for f in ALL_FIELDS})
    st = oracle_lib.track_until(pb, eb, 1, libm=libm)
    assert st == 0
    want = {f: g["final_" + f] for f in ALL_FIELDS}
    assert_particles_equal(particles_dict(p), want, what=name)


def _bbsimple_lattice(g):
    eb = CBuffer()
    for i, t in enumerate(g["types"]):
        t = int(t)
        kw = {k[len("e%d_" % i):]: g[k] for k in g.files if k.startswith("e%d_" % i)}
        if t == 2:
            be.Drift(cbuffer=eb, length=float(kw["length"]))
        elif t == 4:
            be.Multipole(cbuffer=eb, bal=kw["bal"], length=float(kw["length"]),
                         hxl=float(kw["hxl"]), hyl=float(kw["hyl"]))
        elif t == 5:
            be.Cavity(cbuffer=eb, voltage=float(kw["voltage"]),
                      frequency=float(kw["frequency"]), lag=float(kw["lag"]))
        elif t == 8:
            be.BeamBeam4D(cbuffer=eb, size=int(kw["size"]), data=kw["data"])
        elif t == 9:
            be.BeamBeam6D(cbuffer=eb, size=int(kw["size"]), data=kw["data"])
        else:
            raise AssertionError("unexpected type %d" % t)
    return eb


@pytest.mark.parametrize("libm", [True, False])
def test_bbsimple_element_by_element(libm):
    """tests/sixtracklib/common/test_track_c99.cpp:167-290 procedure."""
    g = _load("bbsimple_golden.npz")
    eb = _bbsimple_lattice(g)
    n_el = eb.n_objects
    assert int(g["n_dumps"]) >= n_el + 1
    for ii in range(n_el):
        start = {f: g["p%d_%s" % (ii, f)] for f in ALL_FIELDS}
        want = {f: g["p%d_%s" % (ii + 1, f)] for f in ALL_FIELDS}
        pb, p = particles_from_arrays(start)
        for k in range(p.num_particles):
            assert oracle_lib.track_element(pb, eb, k, ii, libm=libm) == 0
        got = particles_dict(p)
        # the reference compares the real fields only
        # (NS(Particles_compare_real_values_with_treshold)), at 1e-15
        assert_particles_equal(got, want, abs_tol=1e-15, check_index=False,
                               what="element %d (type %d)" % (ii, g["types"][ii]))


@pytest.mark.parametrize("libm", [True, False])
def test_cerrf_faddeeva_grid(libm):
    xyw = _load("faddeeva_golden.npz")["xyw"]
    worst = 0.0
    for x, y, wr, wi in xyw:
        r, i = oracle_lib.cerrf(x, y, libm=libm)
        den = max(abs(complex(wr, wi)), 1e-300)
        worst = max(worst, abs(complex(r - wr, i - wi)) / den)
    assert worst < 4.5e-10, worst


def test_cerrf_against_scipy_wofz():
    """Independent check: scipy.special.wofz is w(z) = exp(-z^2) erfc(-iz)."""
    from scipy.special import wofz
    rng = np.random.default_rng(7)
    z = rng.uniform(-8, 8, 500) + 1j * rng.uniform(-6, 6, 500)
    for zz in z:
        r, i = oracle_lib.cerrf(zz.real, zz.imag, libm=True)
        w = wofz(zz)
        assert abs(complex(r, i) - w) <= 1e-9 * max(abs(w), 1e-3), zz

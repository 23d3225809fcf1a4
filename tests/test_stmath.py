"""Project-owned elementary functions (include/sixtracklib_amd/stmath.h)
against the platform libm (numpy), on the host build of the oracle.

The kernels compile the same header, so these accuracy bounds hold on the
GPU too (the GPU parity tests then check host and device bit for bit).
The reference calls libm directly (be_cavity/track.h:55 sin,
be_beamfields/gauss_fields.h exp, faddeeva_cern.h exp/cos/sin/pow): the
bound below is what separates the two trackers on lattices with cavities
or beam-beam elements.
"""
import math

import numpy as np
import pytest

import oracle_lib


def _ulp_diff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ia = a.view(np.int64).astype(np.int64)
    ib = b.view(np.int64).astype(np.int64)
    # map to a monotone integer line (negative floats reversed)
    ia = np.where(ia < 0, np.int64(-(2**63)) - ia, ia)
    ib = np.where(ib < 0, np.int64(-(2**63)) - ib, ib)
    return np.abs(ia - ib)


def _vec(f, xs):
    return np.array([f(float(x)) for x in xs])


@pytest.fixture(scope="module")
def lib():
    return oracle_lib.oracle()


def test_sin_cos_within_two_ulp_of_libm(lib):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-4.0, 4.0, 3000), rng.uniform(-1e4, 1e4, 3000),
                         # near the quadrant boundaries k*pi/2
                         np.arange(-40, 41) * (np.pi / 2) + rng.uniform(-1e-6, 1e-6, 81)])
    s = _vec(lib.oracle_sin, xs)
    c = _vec(lib.oracle_cos, xs)
    # relative to libm; near zeros of sin/cos compare absolutely
    for got, ref in ((s, np.sin(xs)), (c, np.cos(xs))):
        big = np.abs(ref) > 1e-3
        assert np.max(_ulp_diff(got[big], ref[big])) <= 2
        assert np.max(np.abs(got[~big] - ref[~big])) <= 4e-16


def test_sin_cos_identities_and_quadrants(lib):
    xs = np.linspace(-50.0, 50.0, 4001)
    s = _vec(lib.oracle_sin, xs)
    c = _vec(lib.oracle_cos, xs)
    assert np.max(np.abs(s * s + c * c - 1.0)) < 4e-16
    # odd / even symmetry is exact (reduction and polynomials are symmetric)
    assert np.array_equal(_vec(lib.oracle_sin, -xs), -s)
    assert np.array_equal(_vec(lib.oracle_cos, -xs), c)
    # sin(x + pi/2) = cos(x) within rounding of the shifted argument
    assert np.max(np.abs(_vec(lib.oracle_sin, xs + np.pi / 2) - c)) < 1e-14


def test_sin_cos_special_values(lib):
    assert lib.oracle_sin(0.0) == 0.0
    assert math.copysign(1.0, lib.oracle_sin(-0.0)) == -1.0
    assert lib.oracle_cos(0.0) == 1.0
    for v in (math.inf, -math.inf, math.nan):
        assert math.isnan(lib.oracle_sin(v))
        assert math.isnan(lib.oracle_cos(v))
    # supported domain |x| < 2^30: still within a few ulp of libm near the top
    for v in (1.0e9, -1.0e9 - 0.25, 2.0**29 + 0.5):
        assert abs(lib.oracle_sin(v) - math.sin(v)) < 1e-6
        assert abs(lib.oracle_cos(v) - math.cos(v)) < 1e-6
    # outside it: NaN, with no out-of-range integer conversion on the way
    for v in (2.0**30, 1e22, -3.3e300, 2.0**63, 2.0**80 + 2.0**28):
        assert math.isnan(lib.oracle_sin(v)) and math.isnan(lib.oracle_cos(v))


def test_exp_within_two_ulp_of_libm(lib):
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-745.0, 709.0, 4000), rng.uniform(-1.0, 1.0, 2000),
                         np.array([-744.9, -708.5, 0.0, 1e-300, -1e-300, 709.7])])
    got = _vec(lib.oracle_exp, xs)
    ref = np.exp(xs)
    norm = ref > 2.3e-308                       # normal results: ulp bound
    assert np.max(_ulp_diff(got[norm], ref[norm])) <= 2
    assert np.max(np.abs(got[~norm] - ref[~norm])) <= 1e-323
    assert lib.oracle_exp(0.0) == 1.0
    assert lib.oracle_exp(710.0) == math.inf
    assert lib.oracle_exp(-746.0) == 0.0
    assert lib.oracle_exp(-math.inf) == 0.0
    assert math.isnan(lib.oracle_exp(math.nan))


def test_pow_against_libm(lib):
    rng = np.random.default_rng(3)
    x = rng.uniform(1e-3, 1e3, 3000)
    y = rng.uniform(-20.0, 20.0, 3000)
    got = np.array([lib.oracle_pow(float(a), float(b)) for a, b in zip(x, y)])
    ref = np.power(x, y)
    ok = (ref > 1e-300) & (ref < 1e300)
    # exp(y log x): the log's relative error is amplified by |y log x|
    rel = np.abs(got[ok] - ref[ok]) / ref[ok]
    bound = 4e-16 * (1.0 + np.abs(y[ok] * np.log(x[ok])))
    assert np.all(rel <= bound)
    # integer exponents are exact products (binary exponentiation)
    for a in (0.5, 1.7, -2.25, 3.0):
        for n in range(0, 12):
            want = 1.0
            for _ in range(n):
                want *= a
            assert abs(lib.oracle_pow(a, float(n)) - want) <= 4e-16 * abs(want) * max(1, n)

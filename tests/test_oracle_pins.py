"""Pins for the oracle functions the reference's golden data do not cover.

Each check is something the mathematics or the reference's own tests fix,
chosen so that a plausible slip (a dropped term, a wrong sign or index, a
swapped operand) fails it:

* apertures: the reference's known-answer sets
  (tests/sixtracklib/common/beam_elements/test_be_limit_rect_c99.cpp:102-231,
  test_be_limit_ellipse_c99.cpp:100-269) and the strict global aperture
  (be_limit/track.h:54-88; NaN is lost);
* XYShift, SRotation, DipoleEdge: exact closed forms and invariants;
* Multipole: the complex power sum (an algorithm independent of Horner)
  and the textbook sector-bend limit of the curvature terms;
* Cavity: energy conservation (E' = E + q V sin(phase)) and the
  energy-variable identities rpp = 1/(1+delta), rvv = beta/beta0;
* RFMultipole: reduces to Multipole at zero frequency and phase;
* Gaussian fields: the integral representation of the field of a 2-D
  Gaussian charge (both the round and the Bassetti-Erskine branch);
* q-Gaussian: normalisation to 1; space charge: the bunched map at the
  bunch centre equals the coasting map with the same line density, and the
  interpolated-profile map equals the bunched one on a tabulated Gaussian;
* loops: track_line over segments equals track_until
  (test_track_job_track_line_c99.cpp:237-384), element-by-element records
  are a bijection onto (particle, element, turn) carrying those ids.

All on the stmath build of the oracle (the GPU's bit-level partner) unless
a check says otherwise.  CPU only.
"""
import ctypes as ct
import math

import numpy as np
import pytest
from scipy import integrate

import oracle_lib
from helpers import particles_dict
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd import lattices, beam_elements as be
from sixtracklib_amd.particles import particles_in

EPS = np.finfo(np.float64).eps


def _buf():
    return CBuffer(max_slots=4096, max_objects=64, max_pointers=64)


def _run(eb, turns=1, libm=False, **kw):
    """Track a fresh particle set built from ``kw`` for ``turns`` turns."""
    n = max(np.size(v) for v in kw.values()) if kw else 1
    pb, _ = lattices.make_particles(n, **kw)
    assert oracle_lib.track_until(pb, eb, turns, libm=libm) == 0
    return particles_dict(particles_in(pb)), pb


# ------------------------------------------------------------- apertures

def test_limit_rect_known_answers():
    eb = _buf()
    be.LimitRect(cbuffer=eb, min_x=-1.0, max_x=1.0, min_y=-1.0, max_y=1.0)
    xs = [-1.0, 1.0, 0.0, 0.0, -1.0 - EPS, 1.0 + EPS, 0.0, 0.0]
    ys = [0.0, 0.0, -1.0, 1.0, 0.0, 0.0, -1.0 - EPS, 1.0 + EPS]
    want = [1, 1, 1, 1, 0, 0, 0, 0]
    got, _ = _run(eb, x=np.array(xs), y=np.array(ys))
    assert list(got["state"]) == want
    # lost particles stay lost whatever their position
    pb, _ = lattices.make_particles(8, x=np.zeros(8), y=np.zeros(8), state=0)
    oracle_lib.track_until(pb, eb, 1)
    q = particles_dict(particles_in(pb))
    assert np.all(q["state"] == 0) and np.all(q["at_turn"] == 0)


def test_limit_ellipse_known_answers():
    eb = _buf()
    be.LimitEllipse(cbuffer=eb, a=1.0, b=1.0)
    c30, s30 = math.cos(math.pi / 180.0 * 30.0), math.sin(math.pi / 180.0 * 30.0)
    xs = [-1.0, 1.0, 0.0, 0.0, -1.0 - EPS, 1.0 + EPS, 0.0, 0.0, 0.99 * c30, 1.02 * c30]
    ys = [0.0, 0.0, -1.0, 1.0, 0.0, 0.0, -1.0 - EPS, 1.0 + EPS, 0.99 * s30, 1.02 * s30]
    want = [1, 1, 1, 1, 0, 0, 0, 0, 1, 0]
    got, _ = _run(eb, x=np.array(xs), y=np.array(ys))
    assert list(got["state"]) == want
    # half axes a = 2 (x) and b = 1 (y)
    eb = _buf()
    be.LimitEllipse(cbuffer=eb, a=2.0, b=1.0)
    pts = [((2.0, 0.0), 1), ((0.0, 1.0), 1), ((0.0, 1.5), 0), ((2.0 + 4 * EPS, 0.0), 0),
           ((1.9, 0.3), 1), ((1.2, 0.85), 0)]          # 0.9925 and 1.0825
    got, _ = _run(eb, x=np.array([p[0][0] for p in pts]), y=np.array([p[0][1] for p in pts]))
    assert list(got["state"]) == [p[1] for p in pts]


def test_limit_rect_ellipse_is_intersection():
    # rectangle |x| <= 0.5, |y| <= 2 and the unit circle: a particle survives
    # only inside both (be_limit/track.h:139-162)
    eb = _buf()
    be.LimitRectEllipse(cbuffer=eb, max_x=0.5, max_y=2.0, a=1.0, b=1.0)
    pts = [((0.5, 0.0), 1),              # on the rectangle's edge
           ((0.5 + 2 * EPS, 0.0), 0),    # just outside the rectangle
           ((0.0, 1.0), 1),              # on the circle
           ((0.0, 1.0 + 2 * EPS), 0),    # just outside the circle
           ((0.25, 0.96), 1),            # 0.9841 <= 1
           ((-0.5, 0.9), 0),             # in the rectangle, 1.06 > 1
           ((0.3, 1.5), 0),              # in the rectangle, outside the circle
           ((0.6, 0.1), 0),              # in the circle, outside the rectangle
           ((-0.6, -0.1), 0)]
    xs = np.array([p[0][0] for p in pts])
    ys = np.array([p[0][1] for p in pts])
    got, _ = _run(eb, x=xs, y=ys)
    assert list(got["state"]) == [p[1] for p in pts]


def test_global_aperture_is_strict_and_loses_nan():
    eb = _buf()
    be.Drift(cbuffer=eb, length=0.0)
    one_m = np.nextafter(1.0, 0.0)
    xs = np.array([one_m, 1.0, -one_m, -1.0, 0.0, 0.0, np.nan, np.inf, 0.0])
    ys = np.array([0.0, 0.0, 0.0, 0.0, one_m, -1.0, 0.0, 0.0, np.nan])
    got, _ = _run(eb, x=xs, y=ys)
    assert list(got["state"]) == [1, 0, 1, 0, 1, 0, 0, 0, 0]
    # lost at the drift: at_element stays 0, at_turn 0
    assert list(got["at_turn"]) == [1, 0, 1, 0, 1, 0, 0, 0, 0]


# ------------------------------------------------------------- linear maps

def test_xyshift_srotation_dipedge_closed_forms():
    rng = np.random.default_rng(1)
    x, y = rng.normal(0, 1e-3, 16), rng.normal(0, 1e-3, 16)
    px, py = rng.normal(0, 1e-5, 16), rng.normal(0, 1e-5, 16)
    kw = dict(x=x, y=y, px=px, py=py)
    # XYShift by (dx, dy) then (-dx, -dy): identity; a shift of 0.25 m is exact
    eb = _buf()
    be.XYShift(cbuffer=eb, dx=0.25, dy=-0.125)
    got, _ = _run(eb, **kw)
    assert np.array_equal(got["x"], x - 0.25) and np.array_equal(got["y"], y + 0.125)
    # SRotation by 90 degrees given exactly (cos 0, sin 1): (x, y) -> (y, -x)
    eb = _buf()
    be.SRotation(cbuffer=eb, cos_z=0.0, sin_z=1.0)
    got, _ = _run(eb, **kw)
    assert np.array_equal(got["x"], y) and np.array_equal(got["y"], -x)
    assert np.array_equal(got["px"], py) and np.array_equal(got["py"], -px)
    # any angle: radii kept; +angle then -angle returns the input
    eb = _buf()
    be.SRotation(cbuffer=eb, angle=37.0)
    be.SRotation(cbuffer=eb, angle=-37.0)
    got, _ = _run(eb, **kw)
    np.testing.assert_allclose(got["x"], x, rtol=0, atol=4 * EPS * 1e-3)
    np.testing.assert_allclose(got["py"], py, rtol=0, atol=4 * EPS * 1e-5)
    eb = _buf()
    be.SRotation(cbuffer=eb, angle=37.0)
    got, _ = _run(eb, **kw)
    np.testing.assert_allclose(np.hypot(got["x"], got["y"]), np.hypot(x, y), rtol=4 * EPS)
    # DipoleEdge: px += r21 x, py += r43 y (powers of two: exact)
    eb = _buf()
    be.DipoleEdge(cbuffer=eb, r21=0.5, r43=-0.25)
    got, _ = _run(eb, **kw)
    assert np.array_equal(got["px"], px + 0.5 * x) and np.array_equal(got["py"], py - 0.25 * y)


def test_multipole_is_the_complex_power_sum():
    rng = np.random.default_rng(2)
    knl = rng.normal(0, 1, 6) * np.array([1e-4, 1e-2, 1.0, 10.0, 100.0, 1e3])
    ksl = rng.normal(0, 1, 6) * np.array([1e-4, 1e-2, 1.0, 10.0, 100.0, 1e3])
    x, y = rng.normal(0, 2e-3, 64), rng.normal(0, 2e-3, 64)
    eb = _buf()
    be.Multipole(cbuffer=eb, knl=knl, ksl=ksl)
    got, _ = _run(eb, x=x, y=y)
    z = x + 1j * y
    kick = sum((knl[n] + 1j * ksl[n]) * z ** n / math.factorial(n) for n in range(6))
    scale = np.max(np.abs(kick))
    np.testing.assert_allclose(got["px"], -kick.real, rtol=0, atol=1e-14 * scale)
    np.testing.assert_allclose(got["py"], kick.imag, rtol=0, atol=1e-14 * scale)


def test_multipole_curvature_is_a_thin_sector_bend():
    # knl[0] = hxl = theta, length L: the design orbit is not kicked; an
    # off-momentum particle gets theta*delta (dispersion); an offset particle
    # gets -theta^2 x / L (weak focusing) and a path length change -theta x
    theta, L = 1e-3, 2.0
    eb = _buf()
    be.Multipole(cbuffer=eb, knl=[theta], hxl=theta, length=L)
    got, _ = _run(eb, x=np.array([0.0, 0.0, 1e-3]), delta=np.array([0.0, 1e-4, 0.0]))
    assert got["px"][0] == 0.0
    assert math.isclose(got["px"][1], theta * 1e-4, rel_tol=1e-12)
    assert math.isclose(got["px"][2], -theta * theta * 1e-3 / L, rel_tol=1e-12)
    assert math.isclose(got["zeta"][2], -theta * 1e-3, rel_tol=1e-12)


# ------------------------------------------------------------- energy

@pytest.mark.parametrize("libm", [False, True])
def test_cavity_conserves_energy(libm):
    """lag 90 deg at zeta = 0: the particle gains q V exactly
    (be_cavity/track.h:39-67); the energy variables follow from
    E' = E + q V (particles.h:5255-5297)."""
    p0c, mass0 = 6.5e12, lattices.PROTON_MASS_EV
    V = 5e9                                              # large: delta' ~ 1e-3
    eb = _buf()
    be.Cavity(cbuffer=eb, voltage=V, frequency=400e6, lag=90.0)
    d0 = np.array([0.0, 1e-4, -2e-4])
    q, _ = _run(eb, libm=libm, delta=d0, p0c=p0c, mass0=mass0)
    e0 = np.sqrt((p0c * (1 + d0)) ** 2 + mass0 ** 2)
    e1 = e0 + V
    d1 = np.sqrt(e1 ** 2 - mass0 ** 2) / p0c - 1.0
    np.testing.assert_allclose(q["delta"], d1, rtol=1e-11)
    np.testing.assert_array_equal(q["rpp"], 1.0 / (q["delta"] + 1.0))
    E0 = math.sqrt(p0c ** 2 + mass0 ** 2)
    beta0 = p0c / E0
    beta1 = np.sqrt(e1 ** 2 - mass0 ** 2) / e1
    np.testing.assert_allclose(q["rvv"], beta1 / beta0, rtol=1e-14)
    # psigma = (E - E0) / (beta0^2 E0), E0 the reference energy
    np.testing.assert_allclose(q["psigma"], (e1 - E0) / (beta0 * beta0 * E0), rtol=1e-9)
    assert np.all(q["zeta"] == 0.0)
    # lag 0 at zeta = 0: sin(phase) = 0, the energy variables keep their
    # values up to the rounding of the delta -> ptau -> delta round trip
    eb = _buf()
    be.Cavity(cbuffer=eb, voltage=V, frequency=400e6, lag=0.0)
    pb, _ = lattices.make_particles(3, delta=d0, p0c=p0c, mass0=mass0)
    before = particles_dict(particles_in(pb))
    oracle_lib.track_until(pb, eb, 1, libm=libm)
    after = particles_dict(particles_in(pb))
    np.testing.assert_allclose(after["delta"], before["delta"], rtol=0, atol=4 * EPS)
    np.testing.assert_allclose(after["rvv"], before["rvv"], rtol=4 * EPS)
    np.testing.assert_allclose(after["rpp"], before["rpp"], rtol=4 * EPS)


def test_rf_multipole_reduces_to_multipole():
    rng = np.random.default_rng(4)
    knl, ksl = rng.normal(0, 1, 4) * [1e-3, 1e-2, 1, 10], rng.normal(0, 1, 4) * [1e-3, 1e-2, 1, 10]
    x, y = rng.normal(0, 1e-3, 32), rng.normal(0, 1e-3, 32)
    ea, eb2 = _buf(), _buf()
    be.RFMultipole(cbuffer=ea, voltage=0.0, frequency=0.0, lag=0.0, knl=knl, ksl=ksl)
    be.Multipole(cbuffer=eb2, knl=knl, ksl=ksl)
    a, _ = _run(ea, x=x, y=y)
    b, _ = _run(eb2, x=x, y=y)
    # the RF form sums ascending powers, the plain one uses Horner
    for f in ("px", "py"):
        np.testing.assert_allclose(a[f], b[f], rtol=0, atol=4e-15 * np.max(np.abs(b[f])))
    for f in ("x", "y", "delta", "zeta"):
        np.testing.assert_array_equal(a[f], b[f])


# ------------------------------------------------------------- fields

def _field_integral(x, y, sx, sy):
    """E of a unit 2-D Gaussian line charge, without the 1/(2 pi eps0)
    factor: (x, y) * int_0^inf exp(-x^2/(2sx^2+q) - y^2/(2sy^2+q))
    / ((2sx^2+q) resp. (2sy^2+q)) / sqrt((2sx^2+q)(2sy^2+q)) dq, evaluated
    over u = ln q so that the quadrature sees a smooth, finite range."""
    def f(u, which):
        q = math.exp(u)
        ax, ay = 2 * sx * sx + q, 2 * sy * sy + q
        return q * math.exp(-x * x / ax - y * y / ay) / (ax if which == 0 else ay) \
            / math.sqrt(ax * ay)
    c = math.log(2 * min(sx, sy) ** 2)
    kw = dict(epsabs=0, epsrel=1e-13, limit=1000, points=[c + k for k in range(-10, 30, 2)])
    ix = integrate.quad(f, c - 40, c + 60, args=(0,), **kw)[0]
    iy = integrate.quad(f, c - 40, c + 60, args=(1,), **kw)[0]
    return x * ix, y * iy


@pytest.mark.parametrize("sx,sy", [(1e-3, 1e-3), (2e-3, 1e-3), (1e-3, 3e-3)])
def test_gauss_field_matches_integral_representation(sx, sy):
    """be_beamfields/gauss_fields.h: the round formula for sx == sy, the
    Bassetti-Erskine formula (Faddeeva function, swapped axes for sx < sy)
    otherwise.  Both equal the integral times 1/(2 pi eps0).  The round
    branch is exact to rounding; the elliptic one carries the CERN w(z)
    error (<= 4.3e-10 relative against scipy's wofz), amplified by the
    cancellation between its two w terms when a coordinate is far below
    sigma, so its bound is 1e-9 * max(1, sigma_max / |coordinate|)."""
    lib = oracle_lib.oracle()
    d = ct.c_double
    k = 1.0 / (2.0 * math.pi * 8.854187817620e-12)       # physics_constants.h eps0
    round_beam = sx == sy
    for (x, y) in [(3e-4, 2e-4), (-2e-3, 1e-3), (4e-3, -5e-3), (1e-5, 7e-4)]:
        ex, ey, gx, gy = d(), d(), d(), d()
        lib.oracle_gauss_field(x, y, sx, sy, 1e-10, 1, ct.byref(ex), ct.byref(ey),
                               ct.byref(gx), ct.byref(gy))
        ix, iy = _field_integral(x, y, sx, sy)
        for got, want, coord in ((ex.value, k * ix, x), (ey.value, k * iy, y)):
            tol = 1e-14 if round_beam else 1e-9 * max(1.0, max(sx, sy) / abs(coord))
            assert abs(got / want - 1.0) <= tol, (sx, sy, x, y, got, want)


def test_q_gaussian_is_normalised():
    for q in (0.6, 0.95, 1.0, 1.2, 1.6):
        cq = be.q_gauss_cq(q)
        sb = 1.0 / math.sqrt(2.0)                          # bunch length 1
        if abs(1 - q) < 1e-6:
            f = lambda z: sb / cq * math.exp(-(sb * z) ** 2)
        else:
            f = lambda z, q=q: sb / cq * max(0.0, 1 + (1 - q) * -(sb * z) ** 2) ** (1 / (1 - q))
        total = integrate.quad(f, -np.inf, np.inf, epsabs=0, epsrel=1e-12, limit=400)[0]
        assert math.isclose(total, 1.0, rel_tol=1e-9), (q, total)
    assert be.q_gauss_cq(1.0) == math.sqrt(math.pi)


def test_space_charge_maps_agree_on_one_line_density():
    """Bunched (q-Gaussian, q = 1) at the bunch centre == coasting with the
    same line density N / (sqrt(2 pi) sigma_z); interpolated profile on a
    tabulated Gaussian == bunched at a table node."""
    N, sz, sx, sy = 1e11, 0.1, 1e-3, 2e-3
    x, y = np.array([3e-4, -1e-3, 2e-3]), np.array([2e-4, 5e-4, -3e-3])
    eb_q = _buf()
    be.SCQGaussProfile(cbuffer=eb_q, number_of_particles=N, bunchlength_rms=sz, sigma_x=sx,
                       sigma_y=sy, length=10.0, q_param=1.0, cq=math.sqrt(math.pi))
    eb_c = _buf()
    be.SCCoasting(cbuffer=eb_c, number_of_particles=N, circumference=math.sqrt(2 * math.pi) * sz,
                  sigma_x=sx, sigma_y=sy, length=10.0)
    a, _ = _run(eb_q, x=x, y=y)
    b, _ = _run(eb_c, x=x, y=y)
    assert np.max(np.abs(a["px"])) > 0
    np.testing.assert_allclose(a["px"], b["px"], rtol=1e-14)
    np.testing.assert_allclose(a["py"], b["py"], rtol=1e-14)
    # off-momentum: the bunched kick carries the Lorentz-force factor
    # (1 - beta0 beta) and the transit time L / (beta c): kick ~ (1 - beta0
    # beta) / beta, beta = p / E of the particle; the coasting kick of the
    # reference scales as 1 / rvv = beta0 / beta (be_beamfields/track.h:154-268)
    p0c, m0 = 1e9, lattices.PROTON_MASS_EV                # beta0 ~ 0.73
    dl = np.full(3, 1e-2)
    a0, _ = _run(eb_q, x=x, y=y, p0c=p0c, mass0=m0)
    a1, _ = _run(eb_q, x=x, y=y, p0c=p0c, mass0=m0, delta=dl)
    b0, _ = _run(eb_c, x=x, y=y, p0c=p0c, mass0=m0)
    b1, _ = _run(eb_c, x=x, y=y, p0c=p0c, mass0=m0, delta=dl)
    beta0 = p0c / math.hypot(p0c, m0)
    beta = p0c * 1.01 / math.hypot(p0c * 1.01, m0)
    want = ((1 - beta0 * beta) / beta) / ((1 - beta0 * beta0) / beta0)
    np.testing.assert_allclose(a1["px"] / a0["px"], want, rtol=1e-13)
    np.testing.assert_allclose(b1["py"] / b0["py"], beta0 / beta, rtol=1e-13)
    # interpolated profile, method 0 (be_beamfields.h:3385-3410): the value
    # at node i plus (z - z_i) times the stored derivative at node i.  A
    # Gaussian table whose node 128 sits exactly at z = 0 (dz = 2^-7), with
    # forward-difference derivatives: at z = 0 the kick equals the bunched
    # one; half-way between nodes 128 and 129 it is the mean of the two
    # node values times the same field.
    dz = 2.0 ** -7
    zs = -1.0 + dz * np.arange(257)
    lam = N * np.exp(-zs ** 2 / (2 * sz * sz)) / (math.sqrt(2 * math.pi) * sz)
    slope = np.append(np.diff(lam) / dz, 0.0)
    # the table lives in a buffer of its own, as the reference's stored
    # buffers keep it; the element holds its address
    tb = _buf()
    be.LineDensityProfileData(cbuffer=tb, method=0, capacity=len(zs), z0=zs[0], dz=dz,
                              num_values=len(zs), values=lam, derivatives=slope)
    eb_i = _buf()
    be.SCInterpolatedProfile(cbuffer=eb_i, number_of_particles=1.0, sigma_x=sx,
                             sigma_y=sy, length=10.0,
                             interpol_data_addr=tb.get_object_address(0))
    assert lam[128] == N / (math.sqrt(2 * math.pi) * sz)
    c, _ = _run(eb_i, x=x, y=y)
    np.testing.assert_allclose(c["px"], a["px"], rtol=1e-14)
    np.testing.assert_allclose(c["py"], a["py"], rtol=1e-14)
    h, _ = _run(eb_i, x=x, y=y, zeta=np.full(3, dz / 2))
    ratio = 0.5 * (lam[128] + lam[129]) / lam[128]
    np.testing.assert_allclose(h["px"], c["px"] * ratio, rtol=1e-14)
    # outside the table: no kick
    o, _ = _run(eb_i, x=x, y=y, zeta=np.full(3, 1.5))
    assert np.all(o["px"] == 0.0) and np.all(o["py"] == 0.0)


# ------------------------------------------------------------- loops

def test_track_line_segments_equal_track_until():
    """The reference's LostParticleBehaviour set-up: DriftExact 5 m, quad
    knl[1] = 1e-2 (L = 1), LimitRect +-0.1, DriftExact 5 m; 100 particles at
    x = 0.002 i; 2 turns.  Line tracking over segments with finish_turn on
    the last equals until-turn tracking."""
    eb = _buf()
    be.DriftExact(cbuffer=eb, length=5.0)
    be.Multipole(cbuffer=eb, knl=[0.0, 1e-2], length=1.0)
    be.LimitRect(cbuffer=eb, min_x=-0.1, max_x=0.1, min_y=-0.1, max_y=0.1)
    be.DriftExact(cbuffer=eb, length=5.0)
    x = 0.002 * np.arange(100)
    pa, _ = lattices.make_particles(100, x=x, p0c=10e9, mass0=1e9)
    pl = CBuffer.frombytes(pa.tobytes())
    oracle_lib.track_until(pa, eb, 2)
    for _turn in range(2):
        oracle_lib.track_line(pl, eb, 0, 2, False)
        oracle_lib.track_line(pl, eb, 2, 4, True)
    a, b = particles_dict(particles_in(pa)), particles_dict(particles_in(pl))
    assert np.count_nonzero(a["state"] == 0) > 10
    for f in a:
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)


def _ebe_run(eb, n, turns, monitor_out=None):
    """Element-by-element tracking of n particles over ``turns`` turns into
    a fresh output set; returns (ebe records, final particles)."""
    pb, _ = lattices.make_particles(n, x=np.linspace(1e-4, 5e-4, n), px=1e-6)
    n_el = eb.n_objects
    out, _ = lattices.make_particles(n * n_el * turns, particle_id=-1, state=0)
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=n, num_elements_to_store=n_el,
                               num_turns_to_store=turns, min_particle_id=0, min_element_id=0,
                               min_turn=0, max_particle_id=n - 1, max_element_id=n_el - 1,
                               max_turn=turns - 1, is_rolling=0,
                               out_store_addr=out.get_object_address(0))
    assert oracle_lib.track_elem_by_elem(pb, eb, cfg, turns) == 0
    return particles_dict(particles_in(out)), particles_dict(particles_in(pb))


def test_elem_by_elem_records_are_a_bijection():
    """Store slot k holds (turn, element, particle) = the mixed-radix digits
    of k (output/elem_by_elem_config.h: order 0), each record carrying those
    ids; the record at element e + 1 is the record at e tracked through e."""
    eb = lattices.fodo_ring(n_cells=3)
    n, turns = 5, 3
    n_el = eb.n_objects
    rec, final = _ebe_run(eb, n, turns)
    k = np.arange(n * n_el * turns)
    assert np.array_equal(rec["particle_id"], k % n)
    assert np.array_equal(rec["at_element"], (k // n) % n_el)
    assert np.array_equal(rec["at_turn"], k // (n * n_el))
    assert np.all(rec["state"] == 1)
    # record (t, e+1) == record (t, e) tracked through element e
    for t, e in [(0, 0), (1, 4), (2, n_el - 2)]:
        lo = slice((t * n_el + e) * n, (t * n_el + e + 1) * n)
        cb, _ = lattices.make_particles(n)
        view = particles_in(cb)
        for f in rec:
            type(view).__dict__[f]._view(view)[:] = rec[f][lo]
        for i in range(n):
            assert oracle_lib.track_element(cb, eb, i, e) == 0
        nxt = particles_dict(view)
        hi = slice(lo.start + n, lo.stop + n)
        for f in ("x", "px", "y", "py", "zeta", "delta", "at_element"):
            np.testing.assert_array_equal(nxt[f], rec[f][hi], err_msg=f)
    # last turn's last record tracked through the last element = final state
    assert np.all(final["at_turn"] == turns)


def test_monitor_records_equal_elem_by_elem_records():
    """A turn-ordered BeamMonitor at the end of the ring stores, at each
    turn, exactly the element-by-element record at its own position."""
    eb = lattices.fodo_ring(n_cells=3)
    n, turns = 4, 3
    mon = be.BeamMonitor(cbuffer=eb, num_stores=turns, start=0, skip=1,
                         min_particle_id=0, max_particle_id=n - 1, is_rolling=0,
                         is_turn_ordered=1)
    be.Drift(cbuffer=eb, length=0.5)
    m_idx = eb.n_objects - 2
    mout, _ = lattices.make_particles(n * turns, particle_id=-1, state=0)
    mon.out_address = mout.get_object_address(0)
    rec, _ = _ebe_run(eb, n, turns)
    m = particles_dict(particles_in(mout))
    n_el = eb.n_objects
    for t in range(turns):
        want = slice((t * n_el + m_idx) * n, (t * n_el + m_idx + 1) * n)
        got = slice(t * n, (t + 1) * n)
        for f in m:
            np.testing.assert_array_equal(m[f][got], rec[f][want], err_msg="%s turn %d" % (f, t))
        assert oracle_lib.oracle().oracle_monitor_store_index(
            eb.get_object_address(m_idx), t, n - 1) == t * n + n - 1

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
* RFMultipole: reduces to Multipole at zero frequency and phase, to a
  Cavity at order 0 with zero strengths, and equals the complex-power
  closed form of its transverse kicks and energy gain
  (be_rfmultipole/track.h:97-118) through the energy relation;
* Gaussian fields: the integral representation of the field of a 2-D
  Gaussian charge (both the round and the Bassetti-Erskine branch);
* q-Gaussian: normalisation to 1 and, through the bunched space-charge
  kick's dependence on z, the closed forms q = 2 (Cauchy), q = 0
  (parabola with compact support), q = 0.5 and 1.5
  (internal/math_qgauss.h:114-179); space charge: the bunched map at the
  bunch centre equals the coasting map with the same line density, and the
  interpolated-profile map equals the bunched one on a tabulated Gaussian;
* loops: track_line over segments equals track_until
  (test_track_job_track_line_c99.cpp:237-384), element-by-element records
  are a bijection onto (particle, element, turn) carrying those ids;
* output indices: a rolling BeamMonitor or element-by-element store holds,
  slot for slot, the tail of a non-rolling store of the same tracking
  (be_monitor/track.h:47-112, output/elem_by_elem_config.h:736-743); a
  particle-ordered monitor is the transpose of a turn-ordered one when
  particles and stores are equal in number, and with more particles than
  stores it loses exactly the particles its index pid_off * P + store
  (be_monitor/track.h:105-106) puts past the end of its output;
* the two oracle builds (stmath and libm transcendentals) give the same
  bits on the benchmark workloads' samples.

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
from sixtracklib_amd.cobject import get_object

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


def test_rf_multipole_order0_without_strengths_is_a_cavity():
    """RFMultipole(V, f, lag) of order 0 with knl = ksl = 0 has no kicks and
    gains q V sin(lag - k tau) like Cavity(V, f, lag); only the rounding of
    k tau differs (k zeta / (beta0 rvv) against ((2 pi / c) f) (zeta /
    (beta0 rvv)))."""
    V, f, lag = 3e9, 400e6, 35.0
    zeta = np.array([0.0, 0.05, -0.12, 0.3])
    ea, ec = _buf(), _buf()
    be.RFMultipole(cbuffer=ea, voltage=V, frequency=f, lag=lag, knl=[0.0], ksl=[0.0])
    be.Cavity(cbuffer=ec, voltage=V, frequency=f, lag=lag)
    a, _ = _run(ea, zeta=zeta)
    c, _ = _run(ec, zeta=zeta)
    assert np.all(a["px"] == 0.0) and np.all(a["py"] == 0.0)
    assert np.all(np.abs(c["delta"]) > 1e-5)
    for fld in ("delta", "psigma", "rvv", "rpp", "zeta"):
        np.testing.assert_allclose(a[fld], c[fld], rtol=1e-13, atol=0, err_msg=fld)


def _rf_closed_form(V, f, lag, knl, ksl, pn, ps, x, y, zeta, p0c, mass0):
    """Kicks and energy gain of be_rfmultipole/track.h:39-118 written as
    complex power sums: dpx + i dpy = sum_n (cos(pn_n - k tau) b_n +
    i cos(ps_n - k tau) a_n) z^n, the multipole energy term
    -q k p0c Re sum_n (sin(pn_n - k tau) b_n + i sin(ps_n - k tau) a_n)
    z^(n+1), b_n = knl_n / n!, a_n = ksl_n / n!, z = x + i y, plus the
    cavity term q V sin(lag - k tau); then delta' from E' = E + dE."""
    k = f * 2.0 * math.pi / 299792458.0
    e0 = math.hypot(p0c, mass0)
    beta0 = p0c / e0
    d2r = math.pi / 180.0
    out = []
    for xi, yi, zi in zip(x, y, zeta):
        ktau = k * zi / beta0                              # rvv = 1 at delta = 0
        z = complex(xi, yi)
        kick, dptr = 0j, 0j
        for n in range(len(knl)):
            b, a = knl[n] / math.factorial(n), ksl[n] / math.factorial(n)
            kick += complex(math.cos(pn[n] * d2r - ktau) * b,
                            math.cos(ps[n] * d2r - ktau) * a) * z ** n
            dptr += complex(math.sin(pn[n] * d2r - ktau) * b,
                            math.sin(ps[n] * d2r - ktau) * a) * z ** (n + 1)
        dE = V * math.sin(lag * d2r - ktau) - k * p0c * dptr.real
        e1 = e0 + dE
        d1 = math.sqrt(e1 * e1 - mass0 * mass0) / p0c - 1.0
        out.append((-kick.real, kick.imag, dE, d1))
    return np.array(out)


def test_rf_multipole_matches_complex_closed_form():
    rng = np.random.default_rng(12)
    V, f, lag = 2e9, 400e6, 20.0
    knl = np.array([1e-3, 2e-2, 0.3, 4.0])
    ksl = np.array([-2e-3, 1e-2, -0.5, 3.0])
    pn = np.array([10.0, 40.0, -75.0, 130.0])
    ps = np.array([-20.0, 65.0, 5.0, -150.0])
    x, y = rng.normal(0, 2e-3, 16), rng.normal(0, 2e-3, 16)
    zeta = rng.normal(0, 0.1, 16)
    p0c, m0 = 6.5e12, lattices.PROTON_MASS_EV
    eb = _buf()
    be.RFMultipole(cbuffer=eb, voltage=V, frequency=f, lag=lag, knl=knl, ksl=ksl, pn=pn, ps=ps)
    got, _ = _run(eb, x=x, y=y, zeta=zeta, p0c=p0c, mass0=m0)
    want = _rf_closed_form(V, f, lag, knl, ksl, pn, ps, x, y, zeta, p0c, m0)
    scale = np.max(np.abs(want[:, :2]))
    np.testing.assert_allclose(got["px"], want[:, 0], rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(got["py"], want[:, 1], rtol=0, atol=1e-13 * scale)
    # the energy gain is O(1e9 eV): delta ~ 1e-4, resolved by the energy
    # relation to ~1e-11 relative (as in the Cavity pin)
    assert np.min(np.abs(want[:, 2])) > 1e7
    np.testing.assert_allclose(got["delta"], want[:, 3], rtol=1e-10)
    # the multipole term alone: the same at V = 0
    eb0 = _buf()
    be.RFMultipole(cbuffer=eb0, voltage=0.0, frequency=f, lag=lag, knl=knl, ksl=ksl, pn=pn, ps=ps)
    got0, _ = _run(eb0, x=x, y=y, zeta=zeta, p0c=p0c, mass0=m0)
    want0 = _rf_closed_form(0.0, f, lag, knl, ksl, pn, ps, x, y, zeta, p0c, m0)
    big = np.abs(want0[:, 3]) > 1e-9
    assert np.count_nonzero(big) > 8
    np.testing.assert_allclose(got0["delta"][big], want0[big, 3], rtol=1e-8)


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


@pytest.mark.parametrize("q", [2.0, 1.5, 0.5, 0.0])
def test_q_gaussian_profile_closed_forms(q):
    """The bunched space-charge kick scales with the q-Gaussian line
    density exp_q(-beta z^2) (internal/math_qgauss.h:114-179, beta =
    1 / (2 sigma_z^2), z = zeta / rvv): relative to the kick at z = 0 it is
    1 / (1 + beta z^2) for q = 2 (Cauchy), (1 + beta z^2 / 2)^-2 for q = 1.5,
    max(0, 1 - beta z^2 / 2)^2 for q = 0.5 and max(0, 1 - beta z^2) for
    q = 0 (compact support: no kick beyond it)."""
    sz = 0.1
    beta = 1.0 / (2.0 * sz * sz)
    zs = np.array([0.0, 0.03, 0.08, 0.12, 0.17, 0.5])
    eb = _buf()
    be.SCQGaussProfile(cbuffer=eb, number_of_particles=1e11, bunchlength_rms=sz, sigma_x=1e-3,
                       sigma_y=1.5e-3, length=10.0, q_param=q)
    n = len(zs)
    got, _ = _run(eb, x=np.full(n, 4e-4), y=np.full(n, -2e-4), zeta=zs)
    u = beta * zs * zs
    want = {2.0: 1.0 / (1.0 + u), 1.5: (1.0 + 0.5 * u) ** -2,
            0.5: np.maximum(0.0, 1.0 - 0.5 * u) ** 2, 0.0: np.maximum(0.0, 1.0 - u)}[q]
    assert got["px"][0] != 0.0
    for fld in ("px", "py"):
        ratio = got[fld] / got[fld][0]
        np.testing.assert_allclose(ratio, want, rtol=1e-13, atol=1e-300, err_msg=fld)
    if q < 1.0:
        assert got["px"][-1] == 0.0 and got["py"][-1] == 0.0


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


# ------------------------------------------------------------- output indices

def _monitor_lattice(n_cells, monitors):
    """fodo_ring(n_cells) with ``monitors`` (dicts of BeamMonitor fields)
    inserted before the cavity; returns the buffer and the monitor
    objects' indices."""
    eb = CBuffer(max_slots=8192, max_objects=512, max_pointers=256)
    idx = []
    for c in range(n_cells):
        be.Drift(cbuffer=eb, length=1.0)
        be.Multipole(cbuffer=eb, knl=[0.0, 0.01 if c % 2 == 0 else -0.01])
        be.Drift(cbuffer=eb, length=1.0)
        be.Multipole(cbuffer=eb, knl=[0.0, 0.0, 0.05])
    for m in monitors:
        idx.append(eb.n_objects)
        be.BeamMonitor(cbuffer=eb, **m)
    be.Cavity(cbuffer=eb, voltage=12e6, frequency=400e6, lag=180.0)
    return eb, idx


def _attach_outputs(eb, idx, sizes):
    outs = []
    for k, sz in zip(idx, sizes):
        ob, _ = lattices.make_particles(sz, particle_id=-1, state=0)
        get_object(eb, k).out_address = ob.get_object_address(0)
        outs.append(ob)
    return outs


def test_rolling_monitor_holds_the_tail_of_a_non_rolling_store():
    """Monitors at one position, start 1, skip 2, 11 turns (stored turns
    1, 3, ..., 9: five stores): non-rolling with 5 stores keeps all of
    them; rolling with 3 stores keeps the last three in slots
    (store index mod 3); both orderings."""
    n, turns = 5, 11
    common = dict(start=1, skip=2, min_particle_id=0, max_particle_id=n - 1)
    mons = [dict(num_stores=5, is_rolling=0, is_turn_ordered=1, **common),
            dict(num_stores=3, is_rolling=1, is_turn_ordered=1, **common),
            dict(num_stores=5, is_rolling=0, is_turn_ordered=0, **common),
            dict(num_stores=5, is_rolling=1, is_turn_ordered=0, **common)]
    eb, idx = _monitor_lattice(4, mons)
    outs = _attach_outputs(eb, idx, [5 * n, 3 * n, 5 * n, 5 * n])
    pb, _ = lattices.make_particles(n, x=np.linspace(1e-4, 3e-4, n), y=np.full(n, 1e-4))
    assert oracle_lib.track_until(pb, eb, turns) == 0
    full, roll = particles_dict(particles_in(outs[0])), particles_dict(particles_in(outs[1]))
    # store s of the full monitor is turn 1 + 2 s; the rolling one keeps
    # stores 2, 3, 4 in slots 2, 0, 1
    assert list(full["at_turn"][::n]) == [1, 3, 5, 7, 9]
    # a record carries its monitor's element index; the rest is the particle
    fields = [f for f in full if f != "at_element"]
    outs_d = [full, roll] + [particles_dict(particles_in(o)) for o in outs[2:]]
    for k, d in zip(idx, outs_d):
        assert np.all(d["at_element"] == k)
    for s_full in (2, 3, 4):
        slot = s_full % 3
        for f in fields:
            np.testing.assert_array_equal(roll[f][slot * n:(slot + 1) * n],
                                          full[f][s_full * n:(s_full + 1) * n], err_msg=f)
    # particles x stores equal (5 x 5): particle order is the transpose
    po = outs_d[2]
    for f in fields:
        np.testing.assert_array_equal(po[f].reshape(n, 5), full[f].reshape(5, n).T, err_msg=f)
    # rolling in particle order with 5 stores of 5 stored turns: no wrap
    pr = outs_d[3]
    for f in fields:
        np.testing.assert_array_equal(pr[f], po[f], err_msg=f)


def test_particle_ordered_monitor_loses_the_particles_its_index_overruns():
    """be_monitor/track.h:105-106 indexes a particle-ordered store by
    pid_off * P + store (P particles), not pid_off * num_stores + store.
    With P = 6 particles and 2 stores the output holds 12 records, so
    particle p (store 0 at turn 0) writes record 6 p: particles 0 and 1
    are stored, 2..5 overrun it, and a failed store marks a particle lost
    (be_monitor/track.h:162-165) at the monitor in turn 0."""
    n = 6
    eb, idx = _monitor_lattice(3, [dict(num_stores=2, start=0, skip=1, min_particle_id=0,
                                        max_particle_id=n - 1, is_rolling=0, is_turn_ordered=0)])
    outs = _attach_outputs(eb, idx, [2 * n])
    pb, _ = lattices.make_particles(n, x=np.full(n, 1e-4))
    oracle_lib.track_until(pb, eb, 3)
    fin = particles_dict(particles_in(pb))
    assert list(fin["state"]) == [1, 1, 0, 0, 0, 0]
    assert np.all(fin["at_element"][2:] == idx[0]) and np.all(fin["at_turn"][2:] == 0)
    assert np.all(fin["at_turn"][:2] == 3)
    rec = particles_dict(particles_in(outs[0]))
    # particle 0: records 0 (turn 0) and 1 (turn 1); particle 1: 6 and 7
    assert list(rec["particle_id"]) == [0, 0, -1, -1, -1, -1, 1, 1, -1, -1, -1, -1]
    assert list(rec["at_turn"][[0, 1, 6, 7]]) == [0, 1, 0, 1]


def test_rolling_elem_by_elem_store_holds_the_tail_of_a_full_store():
    """The element-by-element store for 2 turns, rolling, after 5 turns
    holds in turn slot t mod 2 the records of turns 3 and 4
    (output/elem_by_elem_config.h:736-743: the index modulo the store's
    capacity), equal record for record to a 5-turn store of the same
    tracking."""
    eb = lattices.fodo_ring(n_cells=4)
    n, E = 7, eb.n_objects
    pa, _ = lattices.make_particles(n, x=np.linspace(-3e-4, 3e-4, n), y=np.full(n, 2e-4))
    pb = CBuffer.frombytes(pa.tobytes())
    outs = []
    for p, nt, rolling in ((pa, 5, 0), (pb, 2, 1)):
        ocb = CBuffer(max_slots=22 * n * E * nt + 64)
        from sixtracklib_amd.particles import Particles
        Particles(cbuffer=ocb, num_particles=n * E * nt)
        cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=n, num_elements_to_store=E,
                                   num_turns_to_store=nt, min_particle_id=0, min_element_id=0,
                                   min_turn=0, max_particle_id=n - 1, max_element_id=E - 1,
                                   max_turn=nt - 1, is_rolling=rolling,
                                   out_store_addr=ocb.get_object_address(0))
        assert oracle_lib.track_elem_by_elem(p, eb, cfg, 5) == 0
        outs.append(particles_dict(particles_in(ocb)))
    full, roll = outs
    per_turn = n * E
    for t in (3, 4):
        slot = t % 2
        for f in full:
            np.testing.assert_array_equal(roll[f][slot * per_turn:(slot + 1) * per_turn],
                                          full[f][t * per_turn:(t + 1) * per_turn], err_msg=f)


# ------------------------------------------------------------- two builds

def _two_builds(eb, pb, turns):
    a = CBuffer.frombytes(pb.tobytes())
    b = CBuffer.frombytes(pb.tobytes())
    assert oracle_lib.track_until(a, eb, turns, libm=False, threads=8) == 0
    assert oracle_lib.track_until(b, eb, turns, libm=True, threads=8) == 0
    from helpers import assert_particles_equal
    assert_particles_equal(particles_dict(particles_in(a)), particles_dict(particles_in(b)),
                           what="stmath vs libm")


def test_stmath_and_libm_oracles_agree_on_c1():
    """The 402-element FODO ring, 1,000 particles, 100 turns (the complete
    CPU-runnable configuration): the oracle built with stmath (the GPU's
    transcendentals) and with libm (the reference's arithmetic) give the
    same bits, so the exact GPU build, bit-identical to the first, is
    bit-identical to the second here."""
    eb = lattices.fodo_ring()
    pb, _ = lattices.gaussian_beam(1000, seed=0, sigma_x=4e-4, sigma_y=4e-4, sigma_px=4e-6,
                                   sigma_py=4e-6)
    _two_builds(eb, pb, 100)


def test_stmath_and_libm_oracles_agree_on_c2_and_c3_samples():
    """256 particles of the 20,000-element ring's polar grid for 2 turns,
    and 128 matched particles for one turn of the collective ring
    (beam-beam, space charge): the same bits from both builds."""
    eb = lattices.lhc_like_ring(seed=0)
    sig = lattices.twiss_sigma(beta=236.0)
    pb, _ = lattices.polar_grid(4, 64, sigma_x=sig, sigma_y=sig, max_sigma=20.0, delta=2.7e-4)
    _two_builds(eb, pb, 2)
    pb, _ = lattices.matched_gaussian_beam(128, seed=5)
    _two_builds(lattices.collective_ring(), pb, 1)


# ------------------------------------------------------------ symplecticity

S4 = np.array([[0.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 0.0, 0.0],
               [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, -1.0, 0.0]])


def test_drift_multipole_map_is_symplectic():
    """A line of drifts and thin multipoles (normal and skew, orders 0-4)
    at fixed delta is a symplectic map of (x, px, y, py): its Jacobian J
    satisfies J S J^T = S.  The drift moves x by px rpp L (rpp fixed), and
    the multipole kick (-Re f, Im f) with f(x + iy) = sum (knl + i ksl)
    z^n / n! is the gradient of Re F, F' = f, so its Jacobian is symmetric
    (Cauchy-Riemann).  A transposed or mis-signed operand in the complex
    Horner step breaks the symmetry; closed forms of single kicks may not
    see a slip that is symmetric in x and y.  J from Richardson-extrapolated
    central differences of the oracle, at three orbits with delta = 3e-4."""
    rng = np.random.default_rng(41)
    eb = _buf()
    for k in range(12):
        be.Drift(cbuffer=eb, length=float(rng.uniform(0.5, 3.0)))
        order = k % 5
        scale = np.array([1e-4, 5e-2, 1.0, 20.0, 300.0])[:order + 1]
        be.Multipole(cbuffer=eb, knl=rng.normal(0, 1, order + 1) * scale,
                     ksl=rng.normal(0, 1, order + 1) * scale)
    fields = ("x", "px", "y", "py")
    h = np.array([1e-6, 1e-7, 1e-6, 1e-7])
    for orbit in ([1e-3, 2e-5, -5e-4, 1e-5], [-2e-3, -1e-5, 1.5e-3, 3e-5],
                  [4e-4, 4e-5, 2e-4, -4e-5]):
        x0 = np.array(orbit)
        starts = []
        for c in range(4):
            for st in (1.0, -1.0, 0.5, -0.5):
                v = x0.copy()
                v[c] += st * h[c]
                starts.append(v)
        starts = np.array(starts)
        got, _ = _run(eb, **{f: starts[:, k] for k, f in enumerate(fields)},
                      delta=np.full(len(starts), 3e-4))
        end = np.stack([got[f] for f in fields], 1).reshape(4, 4, 4)
        d1 = (end[:, 0] - end[:, 1]) / (2 * h[:, None])
        d2 = (end[:, 2] - end[:, 3]) / h[:, None]
        J = ((4 * d2 - d1) / 3).T                       # [output][input]
        assert np.max(np.abs(J - np.eye(4))) > 1e-2    # a real nonlinear map
        resid = J @ S4 @ J.T - S4
        assert np.max(np.abs(resid)) < 1e-7, (orbit, resid, J)

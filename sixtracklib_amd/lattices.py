"""Synthetic lattices and particle sets for tests and benchmarks.

All generators are seeded and deterministic.  Workloads (sizes from the
benchmark definitions of this project):

* ``fodo_ring``        -- 100 FODO cells [Drift 1 m, thin quad knl[1] =
  +-0.01, Drift 1 m, thin sextupole knl[2] = 0.05] + Cavity(12 MV, 400 MHz,
  lag 180 deg) + LimitRect(+-0.05 m): 402 elements.
* ``lhc_like_ring``    -- 20,000-element ring built from 400 cells of 50
  elements (27 drifts L~U(0.1, 10) m, 20 thin multipoles of order 0..10:
  focusing/defocusing quads sized for a 72 deg FODO phase advance plus
  sextupoles and random orders, all with N(0, 1e-4) relative errors,
  2 XYShift/SRotation with N(0, 1e-4) offsets/angles, 1 LimitEllipse),
  with 8 of the cells carrying a cavity in place of one multipole.
* ``collective_ring`` -- the 20,000-element ring plus 2 BeamBeam6D
  interaction points (5 slices each), 1 BeamBeam4D and 16 SCQGaussProfile
  kicks: 20,019 elements (element-by-element benchmark).
* ``kitchen_sink``     -- short lattice containing every element type the
  trackers support (parity tests).
* particle sets: Gaussian beams, polar amplitude grids.
"""
from __future__ import annotations

import math

import numpy as np

from .cobjects import CBuffer
from .particles import Particles, update_delta_arrays, PROTON_MASS_EV
from . import beam_elements as be


# ---------------------------------------------------------------- particles

def make_particles(n, p0c=6.5e12, mass0=PROTON_MASS_EV, q0=1.0, x=0.0, px=0.0,
                   y=0.0, py=0.0, zeta=0.0, delta=0.0, cbuffer=None,
                   particle_id=None, at_turn=0, at_element=0, state=1):
    """A particle set with the energy variables derived exactly as the C
    tracker's NS(Particles_update_delta) (particles.h:5459-5497)."""
    cb = cbuffer if cbuffer is not None else CBuffer(
        max_slots=22 * int(n) + 64, max_objects=4, max_pointers=24)
    p = Particles(cbuffer=cb, num_particles=int(n))
    energy0 = math.sqrt(p0c * p0c + mass0 * mass0)
    beta0 = p0c / energy0
    gamma0 = energy0 / mass0
    view = lambda name: type(p).__dict__[name]._view(p)
    view("q0")[:] = q0
    view("mass0")[:] = mass0
    view("p0c")[:] = p0c
    view("beta0")[:] = beta0
    view("gamma0")[:] = gamma0
    view("x")[:] = x
    view("px")[:] = px
    view("y")[:] = y
    view("py")[:] = py
    view("zeta")[:] = zeta
    d = np.broadcast_to(np.asarray(delta, dtype=np.float64), (int(n),)).copy()
    rvv, rpp, psigma = update_delta_arrays(d, np.full(int(n), beta0))
    view("delta")[:] = d
    view("rvv")[:] = rvv
    view("rpp")[:] = rpp
    view("psigma")[:] = psigma
    view("chi")[:] = 1.0
    view("charge_ratio")[:] = 1.0
    view("s")[:] = 0.0
    view("particle_id")[:] = np.arange(n) if particle_id is None else particle_id
    view("at_element")[:] = at_element
    view("at_turn")[:] = at_turn
    view("state")[:] = state
    return cb, p


def gaussian_beam(n, seed=0, sigma_x=1e-3, sigma_px=1e-5, sigma_y=1e-3,
                  sigma_py=1e-5, sigma_zeta=0.075, sigma_delta=1e-4, **kw):
    """x, px, y, py, zeta, delta ~ N(0, sigma) drawn in that order from
    numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, sigma_x, n)
    px = rng.normal(0.0, sigma_px, n)
    y = rng.normal(0.0, sigma_y, n)
    py = rng.normal(0.0, sigma_py, n)
    zeta = rng.normal(0.0, sigma_zeta, n)
    delta = rng.normal(0.0, sigma_delta, n)
    return make_particles(n, x=x, px=px, y=y, py=py, zeta=zeta, delta=delta, **kw)


def polar_grid(n_amplitudes, n_angles, sigma_x, sigma_y, max_sigma=20.0,
               delta=2.7e-4, **kw):
    """Particles on a polar grid of normalised amplitude r in (0, max_sigma]
    sigma and angle theta in [0, pi/2): x = r sigma_x cos(theta),
    y = r sigma_y sin(theta), px = py = 0, amplitude-major order (a wave of
    64 lanes holds one amplitude at 64 angles)."""
    r = np.linspace(max_sigma / n_amplitudes, max_sigma, n_amplitudes)
    th = (np.arange(n_angles) + 0.5) * (0.5 * np.pi / n_angles)
    rr, tt = np.meshgrid(r, th, indexing="ij")
    x = (rr * np.cos(tt) * sigma_x).ravel()
    y = (rr * np.sin(tt) * sigma_y).ravel()
    return make_particles(x.size, x=x, y=y, delta=delta, **kw)


# ----------------------------------------------------------------- lattices

def fodo_ring(n_cells=100, k1=0.01, k2=0.05, drift=1.0, voltage=12e6,
              frequency=400e6, lag=180.0, limit=0.05, cbuffer=None):
    eb = cbuffer if cbuffer is not None else CBuffer(
        max_slots=8 * (4 * n_cells + 2) + 64, max_objects=4 * n_cells + 8,
        max_pointers=2 * n_cells + 8)
    for c in range(n_cells):
        be.Drift(cbuffer=eb, length=drift)
        be.Multipole(cbuffer=eb, knl=[0.0, k1 if c % 2 == 0 else -k1])
        be.Drift(cbuffer=eb, length=drift)
        be.Multipole(cbuffer=eb, knl=[0.0, 0.0, k2])
    be.Cavity(cbuffer=eb, voltage=voltage, frequency=frequency, lag=lag)
    be.LimitRect(cbuffer=eb, min_x=-limit, max_x=limit, min_y=-limit, max_y=limit)
    return eb


def lhc_like_ring(seed=0, n_cells=400, tune=80.31, rel_err=1e-4, k2_scale=0.02,
                  r0=0.017, n_cavities=8, voltage=2e6, frequency=400e6,
                  ellipse_a=0.02, ellipse_b=0.015, mean_drift=5.05,
                  error_seed=None, cbuffer=None, return_info=False, after_cell=None):
    """20,000-element synthetic ring (module docstring).

    Each 50-element cell is [QF | 13 drifts, 9 multipoles, 1 shift/rotation
    (shuffled) | QD | 14 drifts, 9 multipoles, 1 shift/rotation (shuffled) |
    LimitEllipse].  Drift lengths are drawn U(0.1, 10) m and rescaled so
    every half cell has the same total length (13 and 14 x ``mean_drift``):
    the linear optics of all cells are then identical, which keeps the ring
    stable, while the element sequence stays random.  The thin quads give a
    thin-lens FODO phase advance mu with cos(mu) = 1 - l1 l2 k^2 / 2 and
    ``tune`` = n_cells mu / 2pi (non-integer).  Non-quad multipoles have a
    random order 0..10, half of those of order >= 2 carry a design
    sextupole +-``k2_scale``; every multipole gets N(0, 1e-4) relative
    errors on all its coefficients (normalised to the quad strength at the
    reference radius ``r0``).  ``seed`` fixes the structure, ``error_seed``
    (default ``seed``) the error table.  8 cells carry a cavity in place of
    one multipole.  ``after_cell(eb, c)``, when given, may append elements
    after cell ``c``."""
    if error_seed is None:
        error_seed = seed
    rng = np.random.default_rng(seed)
    erng = np.random.default_rng(10_000 + int(error_seed))
    cell_len = 50
    n_elem = n_cells * cell_len
    eb = cbuffer if cbuffer is not None else CBuffer(
        max_slots=n_elem * 16, max_objects=n_elem + 16, max_pointers=n_elem)
    cav_cells = set(np.linspace(0, n_cells - 1, n_cavities).astype(int).tolist())
    l1, l2 = 13 * mean_drift, 14 * mean_drift
    mu = 2.0 * math.pi * tune / n_cells
    k1 = 2.0 * math.sin(mu / 2.0) / math.sqrt(l1 * l2)
    counts = {"drift": 0, "multipole": 0, "xyshift": 0, "srotation": 0,
              "ellipse": 0, "cavity": 0}
    orders = []
    mp_design = []            # (object index, knl design, ksl design, error scale)
    misc_toggle = 0
    for c in range(n_cells):
        halves = []
        for nd, total in ((13, l1), (14, l2)):
            lengths = rng.uniform(0.1, 10.0, nd)
            lengths *= total / lengths.sum()
            slots = ["drift"] * nd + ["mp"] * 9 + ["misc"]
            rng.shuffle(slots)
            halves.append((slots, iter(lengths)))
        cav_done = c not in cav_cells
        for part, (slots, lengths) in zip(("QF", "QD"), halves):
            for s in [part] + slots:
                if s == "drift":
                    be.Drift(cbuffer=eb, length=float(next(lengths)))
                    counts["drift"] += 1
                elif s == "mp" and not cav_done:
                    be.Cavity(cbuffer=eb, voltage=voltage, frequency=frequency, lag=180.0)
                    counts["cavity"] += 1
                    cav_done = True
                elif s in ("QF", "QD", "mp"):
                    order = int(rng.integers(0, 11)) if s == "mp" else int(rng.integers(1, 11))
                    knl = np.zeros(order + 1)
                    ksl = np.zeros(order + 1)
                    if s == "QF":
                        knl[1] = +k1
                    elif s == "QD":
                        knl[1] = -k1
                    elif order >= 2 and rng.random() < 0.5:
                        knl[2] = k2_scale * (1.0 if rng.random() < 0.5 else -1.0)
                    nn = np.arange(order + 1)
                    scale = rel_err * k1 * np.array([math.factorial(int(m)) for m in nn]) \
                        * r0 ** (1.0 - nn)
                    mp_design.append((eb.n_objects, knl.copy(), ksl.copy(), scale))
                    knl = knl + scale * erng.normal(0.0, 1.0, order + 1)
                    ksl = ksl + scale * erng.normal(0.0, 1.0, order + 1)
                    be.Multipole(cbuffer=eb, knl=knl, ksl=ksl)
                    orders.append(order)
                    counts["multipole"] += 1
                elif s == "misc":
                    if misc_toggle == 0:
                        be.XYShift(cbuffer=eb, dx=rng.normal(0, 1e-4), dy=rng.normal(0, 1e-4))
                        counts["xyshift"] += 1
                    else:
                        ang = rng.normal(0, 1e-4)
                        be.SRotation(cbuffer=eb, cos_z=math.cos(ang), sin_z=math.sin(ang))
                        counts["srotation"] += 1
                    misc_toggle ^= 1
        be.LimitEllipse(cbuffer=eb, a=ellipse_a, b=ellipse_b)
        counts["ellipse"] += 1
        if after_cell is not None:
            after_cell(eb, c)
    if return_info:
        return eb, {"counts": counts, "orders": orders, "k1": k1, "mu": mu,
                    "half_cells": (l1, l2), "mp_design": mp_design}
    return eb


def lhc_like_ring_family(n_seeds, seed=0, first_error_seed=0, **kw):
    """``n_seeds`` machine seeds of one ring: the structure of
    lhc_like_ring(seed) with the multipole-error tables of error_seed =
    first_error_seed + k.  Ring k equals lhc_like_ring(seed, error_seed=
    first_error_seed + k) word for word; the structure is generated once and
    only the bal arrays are rewritten (same draws, same arithmetic)."""
    base, info = lhc_like_ring(seed=seed, error_seed=first_error_seed, return_info=True, **kw)
    mp = info["mp_design"]
    # word offset of every multipole's bal array in the buffer
    words = base.words
    bal_off = []
    for idx, knl, _, _ in mp:
        o = base.get_object_word_offset(idx)
        bal_off.append((int(words[o + 4]) - base.pointer) // 8)
    inv_f = [np.array([be._inv_factorial(i) for i in range(len(knl))]) for _, knl, _, _ in mp]
    rings = [base]
    for k in range(1, n_seeds):
        erng = np.random.default_rng(10_000 + int(first_error_seed + k))
        cb = CBuffer.frombytes(base.tobytes())
        w = cb.words.view(np.float64)
        for (idx, knl, ksl, scale), off, inv in zip(mp, bal_off, inv_f):
            n = len(knl)
            nknl = knl + scale * erng.normal(0.0, 1.0, n)
            nksl = ksl + scale * erng.normal(0.0, 1.0, n)
            w[off:off + 2 * n:2] = nknl * inv
            w[off + 1:off + 2 * n:2] = nksl * inv
        rings.append(cb)
    return rings


def adjoint_sublattice(seed=0, n_cells=20, tune=4.31, rel_err=1e-4, k2_scale=0.02,
                       r0=0.017, voltage=2e6, frequency=400e6, mean_drift=5.05,
                       cbuffer=None):
    """1,000-element sub-lattice for differentiable tracking: Drift,
    Multipole of order 1-3 and one Cavity.

    Each 50-element cell is [QF | 12 drifts, 12 multipoles (shuffled) | QD |
    12 drifts, 12 multipoles (shuffled)], drift lengths U(0.1, 10) m
    rescaled so each half cell is 12 x ``mean_drift`` long (identical linear
    optics in every cell, thin-lens FODO phase advance 2 pi tune / n_cells).
    Multipoles have a random order 1..3; half of those of order >= 2 carry a
    design sextupole +-``k2_scale``; all get N(0, ``rel_err``) relative
    errors on every coefficient (normalised to the quad strength at radius
    ``r0``).  Cell 0 has a cavity in place of its first multipole."""
    rng = np.random.default_rng(seed)
    erng = np.random.default_rng(20_000 + int(seed))
    eb = cbuffer if cbuffer is not None else CBuffer(
        max_slots=n_cells * 50 * 16, max_objects=n_cells * 50 + 16, max_pointers=n_cells * 50)
    half = 12 * mean_drift
    mu = 2.0 * math.pi * tune / n_cells
    k1 = 2.0 * math.sin(mu / 2.0) / half
    for c in range(n_cells):
        cav_done = c != 0
        for quad in (+k1, -k1):
            lengths = rng.uniform(0.1, 10.0, 12)
            lengths *= half / lengths.sum()
            lengths = iter(lengths)
            slots = ["drift"] * 12 + ["mp"] * 12
            rng.shuffle(slots)
            for s in ["quad"] + slots:
                if s == "drift":
                    be.Drift(cbuffer=eb, length=float(next(lengths)))
                    continue
                if s == "mp" and not cav_done:
                    be.Cavity(cbuffer=eb, voltage=voltage, frequency=frequency, lag=180.0)
                    cav_done = True
                    continue
                order = 1 if s == "quad" else int(rng.integers(1, 4))
                knl = np.zeros(order + 1)
                ksl = np.zeros(order + 1)
                if s == "quad":
                    knl[1] = quad
                elif order >= 2 and rng.random() < 0.5:
                    knl[2] = k2_scale * (1.0 if rng.random() < 0.5 else -1.0)
                nn = np.arange(order + 1)
                scale = rel_err * k1 * np.array([math.factorial(int(m)) for m in nn]) \
                    * r0 ** (1.0 - nn)
                knl = knl + scale * erng.normal(0.0, 1.0, order + 1)
                ksl = ksl + scale * erng.normal(0.0, 1.0, order + 1)
                be.Multipole(cbuffer=eb, knl=knl, ksl=ksl)
    return eb


def gaussian_slice_centroids(n_slices, sigma_z):
    """Longitudinal centroids of ``n_slices`` equal-charge slices of a
    Gaussian bunch of rms length ``sigma_z``, tail first: slice k spans the
    quantiles k / n .. (k + 1) / n, and its centroid is
    sigma_z * n * (pdf(a) - pdf(b)) for its edges a < b in units of sigma
    (the mean of a Gaussian restricted to [a, b]).  Two slices give
    -+ sigma_z * sqrt(2 / pi), one gives 0."""
    from statistics import NormalDist
    nd = NormalDist()
    edges = [nd.inv_cdf(k / n_slices) if 0 < k < n_slices else None
             for k in range(n_slices + 1)]
    cent = []
    for k in range(n_slices):
        pa = nd.pdf(edges[k]) if edges[k] is not None else 0.0
        pb = nd.pdf(edges[k + 1]) if edges[k + 1] is not None else 0.0
        cent.append(sigma_z * n_slices * (pa - pb))
    return cent


def collective_ring(seed=0, bunch_charge=1.15e11, sigma_ip=16.7e-6, beta_ip=0.55,
                    sigma_z=0.075, phi=142.5e-6, n_slices=5, n_sc=16, beta_arc=236.0,
                    sc_sigma_ratio=0.55, n_cells=400):
    """The 20,000-element ring with collective elements (20,019 elements;
    ``n_cells`` < 400 gives a shorter ring of the same make-up):

    * BeamBeam6D after cells 0 and n_cells / 2 (horizontal and vertical crossing,
      half angle ``phi``): a round Gaussian opposing bunch of
      ``bunch_charge`` charges, sigma_x = sigma_y = ``sigma_ip``,
      sigma_px = sigma_py = sigma_ip / beta_ip, cut into ``n_slices``
      equal-charge slices at the slice centroids of a Gaussian of rms
      ``sigma_z``;
    * BeamBeam4D after cell n_cells / 4 with the same charge and sizes;
    * ``n_sc`` SCQGaussProfile kicks (q = 1, bunch length ``sigma_z``),
      one after every (n_cells / n_sc)-th cell, each covering 1 / n_sc of the
      circumference, with the arc beam size sigma = twiss_sigma(beta_arc)
      horizontally and ``sc_sigma_ratio`` times it vertically (an
      elliptical beam, so the kicks use the Faddeeva-function field).
    """
    cent = gaussian_slice_centroids(n_slices, sigma_z)
    sig_ip2 = sigma_ip * sigma_ip
    sig_p2 = (sigma_ip / beta_ip) ** 2
    sig_arc = twiss_sigma(beta=beta_arc)
    circumference = n_cells * 27 * 5.05
    every = max(1, n_cells // n_sc)

    def bb6d(eb, alpha):
        be.BeamBeam6D(cbuffer=eb, phi=phi, alpha=alpha, x_bb_co=0.0, y_bb_co=0.0,
                      charge_slices=[bunch_charge / n_slices] * n_slices,
                      zeta_slices=cent, sigma_11=sig_ip2, sigma_12=0.0, sigma_13=0.0,
                      sigma_14=0.0, sigma_22=sig_p2, sigma_23=0.0, sigma_24=0.0,
                      sigma_33=sig_ip2, sigma_34=0.0, sigma_44=sig_p2, x_co=0.0,
                      px_co=0.0, y_co=0.0, py_co=0.0, zeta_co=0.0, delta_co=0.0,
                      min_sigma_diff=1e-10, threshold_singular=1e-28, d_x=0.0, d_px=0.0,
                      d_y=0.0, d_py=0.0, d_zeta=0.0, d_delta=0.0, enabled=1)

    def extra(eb, c):
        if c == 0:
            bb6d(eb, 0.0)
        elif c == n_cells // 2:
            bb6d(eb, 0.5 * math.pi)
        elif c == n_cells // 4:
            be.BeamBeam4D(cbuffer=eb, charge=bunch_charge, sigma_x=sigma_ip, sigma_y=sigma_ip,
                          beta_r=1.0, min_sigma_diff=1e-10, x_bb=0.0, y_bb=0.0, d_px=0.0,
                          d_py=0.0, enabled=1.0)
        if c % every == every // 2 and c // every < n_sc:
            be.SCQGaussProfile(cbuffer=eb, number_of_particles=bunch_charge,
                               bunchlength_rms=sigma_z, sigma_x=sig_arc,
                               sigma_y=sc_sigma_ratio * sig_arc,
                               length=circumference / n_sc, q_param=1.0)

    n_elem = 50 * n_cells
    eb = CBuffer(max_slots=n_elem * 16 + 4096, max_objects=n_elem + 64, max_pointers=n_elem + 64)
    # same phase advance per cell as the full ring
    return lhc_like_ring(seed=seed, n_cells=n_cells, tune=80.31 * n_cells / 400.0,
                         n_cavities=min(8, n_cells), cbuffer=eb, after_cell=extra)


def matched_gaussian_beam(n, seed=0, beta=236.0, sigma_zeta=0.075, sigma_delta=1.1e-4):
    """6D Gaussian beam matched to the ring start (alpha = 0, beta
    ``beta``): sigma_x = sigma_y = twiss_sigma(beta), sigma_px = sigma_x /
    beta; longitudinal rms ``sigma_zeta`` and ``sigma_delta``."""
    s = twiss_sigma(beta=beta)
    return gaussian_beam(n, seed=seed, sigma_x=s, sigma_px=s / beta, sigma_y=s,
                         sigma_py=s / beta, sigma_zeta=sigma_zeta, sigma_delta=sigma_delta)


def batch_workload(n_seeds=64, seed=0, min_particles=1000, max_particles=100000,
                   count_seed=1, beta=236.0, max_sigma=20.0, delta=2.7e-4):
    """Multi-seed workload: lhc_like_ring_family(n_seeds, seed) and, per
    seed, U(min_particles, max_particles) particles (numpy
    default_rng(count_seed).integers) on the first points of an
    amplitude-major polar grid (64 angles, amplitudes up to ``max_sigma``
    sigma of twiss_sigma(beta)).  Returns (lattices, particle buffers,
    counts)."""
    rings = lhc_like_ring_family(n_seeds, seed=seed)
    counts = np.random.default_rng(count_seed).integers(min_particles, max_particles + 1,
                                                        n_seeds)
    sig = twiss_sigma(beta=beta)
    parts = []
    for n in counts:
        n = int(n)
        n_amp = (n + 63) // 64
        r = np.linspace(max_sigma / n_amp, max_sigma, n_amp)
        th = (np.arange(64) + 0.5) * (0.5 * np.pi / 64)
        rr, tt = np.meshgrid(r, th, indexing="ij")
        x = (rr * np.cos(tt) * sig).ravel()[:n]
        y = (rr * np.sin(tt) * sig).ravel()[:n]
        pb, _ = make_particles(n, x=x, y=y, delta=delta)
        parts.append(pb)
    return rings, parts, counts


def flops_per_particle_turn(ebuf):
    """Algorithmic FP64 operation count of one particle-turn through
    ``ebuf`` (+ - * / sqrt = 1, transcendental calls counted separately):
    Drift 15 + 2 aperture compares, Multipole 8*order + 4 (+ curvature),
    XYShift 2, SRotation 12, LimitEllipse 5, Cavity 37 + 1 sin,
    DriftExact 17 + 2, LimitRect 0, DipoleEdge 4."""
    from .cobject import get_object
    fl = 0
    trans = 0
    for i in range(ebuf.n_objects):
        t = ebuf.get_object_typeid(i)
        if t == 2:
            fl += 17
        elif t == 3:
            fl += 19
        elif t == 4:
            o = get_object(ebuf, i)
            fl += 8 * int(o.order) + 4
            if o.hxl != 0.0 or o.hyl != 0.0:
                fl += 12
        elif t == 5:
            fl += 37
            trans += 1
        elif t == 6:
            fl += 2
        elif t == 7:
            fl += 12
        elif t == 12:
            fl += 5
        elif t == 16:
            fl += 7
        elif t == 24:
            fl += 4
    return fl, trans


def twiss_sigma(beta=100.0, emit_norm=2.5e-6, p0c=6.5e12, mass0=PROTON_MASS_EV):
    """rms beam size sqrt(beta * eps_n / (beta_rel gamma_rel))."""
    bg = p0c / mass0
    return math.sqrt(beta * emit_norm / bg)


def kitchen_sink(seed=0, cbuffer=None, with_monitor=False):
    """Every supported element type once or more, moderate strengths."""
    rng = np.random.default_rng(seed)
    eb = cbuffer if cbuffer is not None else CBuffer(max_slots=4096, max_objects=128,
                                                     max_pointers=64)
    be.Drift(cbuffer=eb, length=0.5)
    be.Multipole(cbuffer=eb, knl=[1e-5, 0.02, 0.1, 1.0], ksl=[0.0, 1e-4, 0.0, 0.5])
    be.DriftExact(cbuffer=eb, length=0.7)
    be.Multipole(cbuffer=eb, knl=[0.001, -0.01], hxl=0.001, hyl=0.0005, length=1.5)
    be.XYShift(cbuffer=eb, dx=1e-4, dy=-2e-4)
    be.SRotation(cbuffer=eb, angle=3.0)
    be.DipoleEdge(cbuffer=eb, r21=0.002, r43=-0.003)
    be.Cavity(cbuffer=eb, voltage=5e6, frequency=200e6, lag=170.0)
    be.RFMultipole(cbuffer=eb, voltage=1e5, frequency=400e6, lag=10.0,
                   knl=[1e-6, 1e-3, 0.01], ksl=[0.0, 2e-4, 0.0], pn=[0.0, 30.0, 60.0],
                   ps=[10.0, 0.0, -40.0])
    be.LimitRect(cbuffer=eb, min_x=-0.02, max_x=0.02, min_y=-0.015, max_y=0.015)
    be.Drift(cbuffer=eb, length=1.0)
    be.BeamBeam4D(cbuffer=eb, charge=1e11, sigma_x=2e-3, sigma_y=1.5e-3, beta_r=1.0,
                  min_sigma_diff=1e-10, x_bb=1e-4, y_bb=-1e-4, d_px=0.0, d_py=0.0,
                  enabled=1.0)
    be.BeamBeam4D(cbuffer=eb, charge=1e11, sigma_x=2e-3, sigma_y=2e-3, beta_r=1.0,
                  min_sigma_diff=1e-10, x_bb=0.0, y_bb=0.0, d_px=1e-7, d_py=0.0,
                  enabled=1.0)
    be.BeamBeam6D(cbuffer=eb, phi=0.01, alpha=0.3, x_bb_co=1e-4, y_bb_co=0.0,
                  charge_slices=[3e10, 4e10, 3e10], zeta_slices=[-0.05, 0.0, 0.05],
                  sigma_11=4e-6, sigma_12=1e-8, sigma_13=0.0, sigma_14=0.0,
                  sigma_22=1e-9, sigma_23=0.0, sigma_24=0.0, sigma_33=2.5e-6,
                  sigma_34=-1e-8, sigma_44=1e-9, x_co=0.0, px_co=0.0, y_co=0.0,
                  py_co=0.0, zeta_co=0.0, delta_co=0.0, min_sigma_diff=1e-10,
                  threshold_singular=1e-28, d_x=0.0, d_px=0.0, d_y=0.0, d_py=0.0,
                  d_zeta=0.0, d_delta=0.0, enabled=1)
    be.LimitEllipse(cbuffer=eb, a=0.02, b=0.02)
    be.SCCoasting(cbuffer=eb, number_of_particles=1e11, circumference=1000.0,
                  sigma_x=1.2e-3, sigma_y=0.9e-3, length=1.0, x_co=0.0, y_co=0.0)
    be.SCQGaussProfile(cbuffer=eb, number_of_particles=1e11, bunchlength_rms=0.1,
                       sigma_x=1.1e-3, sigma_y=1.1e-3, length=1.0, q_param=1.2)
    be.SCQGaussProfile(cbuffer=eb, number_of_particles=1e11, bunchlength_rms=0.1,
                       sigma_x=1.3e-3, sigma_y=0.8e-3, length=1.0, q_param=1.0)
    be.SCInterpolatedProfile(cbuffer=eb, number_of_particles=1e11, sigma_x=1e-3,
                             sigma_y=1.2e-3, length=1.0, line_density_prof_fallback=0.5)
    be.LimitRectEllipse(cbuffer=eb, max_x=0.02, max_y=0.02, a_squ=4e-4, b_squ=3e-4)
    # orders above the unrolled 0-10 (the generic Horner loop), as in the
    # LHC lattices of the reference's test data (orders up to 14); every
    # order kicks by about 1e-7 at 1 cm, so that each coefficient shows in
    # the result bits
    scale = 1e-7 * np.array([math.factorial(k) / 0.01 ** k for k in range(15)])
    be.Multipole(cbuffer=eb, knl=rng.normal(0.0, 1.0, 12) * scale[:12],
                 ksl=rng.normal(0.0, 1.0, 12) * scale[:12])
    be.Multipole(cbuffer=eb, knl=rng.normal(0.0, 1.0, 15) * scale,
                 ksl=rng.normal(0.0, 1.0, 15) * scale)
    if with_monitor:
        be.BeamMonitor(cbuffer=eb, num_stores=4, start=0, skip=1, is_rolling=1)
    be.Drift(cbuffer=eb, length=0.3)
    return eb

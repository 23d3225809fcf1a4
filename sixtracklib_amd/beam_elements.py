"""Beam-element classes stored in a CObjects buffer.

Mirror of python/sixtracklib/beam_elements.py: same class names, field
names, field order, type ids and constructor conveniences (knl/ksl -> bal,
SRotation(angle), LimitEllipse(a, b), DipoleEdge(h, e1, hgap, fint),
BeamBeam4D(x_bb=...), SCQGaussProfile(q_param) with cq, ``Elements``).
The POD layouts match the C structs cited on each class, so a lattice
built here is byte-compatible with the reference C library and is what the
lattice packer of this package reads.
"""
from __future__ import annotations

import math

import numpy as np

from .cobject import CObject, CField, get_object
from .cobjects import CBuffer

QELEM_SCIPY = 1.602176634e-19   # scipy.constants.e, used by beam_elements.py:11


def _inv_factorial(n):
    return 1.0 / math.factorial(n)


class Drift(CObject):
    """be_drift/be_drift.h:28-31"""
    _typeid = 2
    length = CField(0, "real", default=0.0)


class DriftExact(CObject):
    """be_drift/be_drift.h:159-162"""
    _typeid = 3
    length = CField(0, "real", default=0.0)


class Multipole(CObject):
    """be_multipole/be_multipole.h:27-34; knl/ksl -> bal as
    python/sixtracklib/beam_elements.py:67-108 (bal[2n] = knl[n] * (1/n!))."""
    _typeid = 4
    order = CField(0, "int64", default=0, const=True)
    length = CField(1, "real", default=0.0)
    hxl = CField(2, "real", default=0.0)
    hyl = CField(3, "real", default=0.0)
    bal = CField(4, "real", default=0.0, pointer=True,
                 length="2 * order + 2")

    def __init__(self, order=None, knl=None, ksl=None, bal=None, **kwargs):
        if bal is None and (knl is not None or ksl is not None or
                            order is not None):
            knl = [] if knl is None else list(knl)
            ksl = [] if ksl is None else list(ksl)
            order = 0 if order is None else order
            n = max(order + 1, max(len(knl), len(ksl)))
            nknl = np.zeros(n)
            nknl[:len(knl)] = knl
            nksl = np.zeros(n)
            nksl[:len(ksl)] = ksl
            order = n - 1
            bal = np.zeros(2 * order + 2)
            inv_f = np.array([_inv_factorial(i) for i in range(n)])
            bal[0::2] = nknl * inv_f
            bal[1::2] = nksl * inv_f
            kwargs["bal"] = bal
            kwargs["order"] = order
        elif bal is not None and len(bal) > 0:
            kwargs["bal"] = bal
            kwargs["order"] = (len(bal) - 2) // 2
        super().__init__(**kwargs)

    @property
    def knl(self):
        idx = np.arange(0, len(self.bal), 2)
        return self.bal[idx] * np.array([math.factorial(i // 2) for i in idx])

    @property
    def ksl(self):
        idx = np.arange(0, len(self.bal), 2)
        return self.bal[idx + 1] * np.array([math.factorial(i // 2) for i in idx])

    def set_knl(self, value, order):
        assert order <= self.order
        self.bal[order * 2] = value / math.factorial(order)

    def set_ksl(self, value, order):
        assert order <= self.order
        self.bal[order * 2 + 1] = value / math.factorial(order)


class RFMultipole(CObject):
    """be_rfmultipole/be_rfmultipole.h (order, voltage, frequency, lag,
    bal, phase); python/sixtracklib/beam_elements.py:131-266."""
    _typeid = 256
    order = CField(0, "int64", default=0, const=True)
    voltage = CField(1, "real", default=0.0)
    frequency = CField(2, "real", default=0.0)
    lag = CField(3, "real", default=0.0)
    bal = CField(4, "real", default=0.0, pointer=True, length="2 * order + 2")
    phase = CField(5, "real", default=0.0, pointer=True,
                   length="2 * order + 2")

    def __init__(self, order=None, knl=None, ksl=None, pn=None, ps=None,
                 bal=None, phase=None, **kwargs):
        if bal is None and (knl is not None or ksl is not None or
                            order is not None):
            knl = [] if knl is None else list(knl)
            ksl = [] if ksl is None else list(ksl)
            pn = [] if pn is None else list(pn)
            ps = [] if ps is None else list(ps)
            order = 0 if order is None else order
            n = max(order + 1, len(knl), len(ksl), len(pn), len(ps))
            order = n - 1
            arr = lambda v: np.pad(np.asarray(v, dtype=float), (0, n - len(v)))
            inv_f = np.array([_inv_factorial(i) for i in range(n)])
            bal = np.zeros(2 * n)
            bal[0::2] = arr(knl) * inv_f
            bal[1::2] = arr(ksl) * inv_f
            phase = np.zeros(2 * n)
            phase[0::2] = arr(pn)
            phase[1::2] = arr(ps)
            kwargs.update(bal=bal, phase=phase, order=order)
        elif bal is not None:
            kwargs.update(bal=bal, order=(len(bal) - 2) // 2)
            if phase is not None:
                kwargs["phase"] = phase
        super().__init__(**kwargs)


class Cavity(CObject):
    """be_cavity/be_cavity.h:27-32"""
    _typeid = 5
    voltage = CField(0, "real", default=0.0)
    frequency = CField(1, "real", default=0.0)
    lag = CField(2, "real", default=0.0)


class XYShift(CObject):
    """be_xyshift/be_xyshift.h:26-30"""
    _typeid = 6
    dx = CField(0, "real", default=0.0)
    dy = CField(1, "real", default=0.0)


class SRotation(CObject):
    """be_srotation/be_srotation.h:26-30; beam_elements.py:282-299"""
    _typeid = 7
    cos_z = CField(0, "real", default=1.0)
    sin_z = CField(1, "real", default=0.0)

    def __init__(self, angle=0, **nargs):
        if "cos_z" not in nargs and "sin_z" not in nargs:
            anglerad = angle / 180 * np.pi
            nargs["cos_z"] = np.cos(anglerad)
            nargs["sin_z"] = np.sin(anglerad)
        super().__init__(**nargs)

    @property
    def angle(self):
        return np.arctan2(self.sin_z, self.cos_z)

    @property
    def angle_deg(self):
        return self.angle * (180.0 / np.pi)


class BeamMonitor(CObject):
    """be_monitor/be_monitor.h:35-45"""
    _typeid = 10
    num_stores = CField(0, "int64", default=0)
    start = CField(1, "int64", default=0)
    skip = CField(2, "int64", default=1)
    out_address = CField(3, "uint64", default=0)
    max_particle_id = CField(4, "int64", default=0)
    min_particle_id = CField(5, "int64", default=0)
    is_rolling = CField(6, "int64", default=0)
    is_turn_ordered = CField(7, "int64", default=1)


def append_beam_monitors_to_lattice(beam_elements_buffer,
                                    until_turn_elem_by_elem,
                                    until_turn_turn_by_turn, until_turn,
                                    skip_turns=1, min_particle_id=0,
                                    max_particle_id=0, initial_at_turn=0):
    """python/sixtracklib/beam_elements.py:314-367"""
    added = 0
    start_tbt = max(initial_at_turn, until_turn_elem_by_elem)
    if until_turn_turn_by_turn > start_tbt:
        BeamMonitor(start=start_tbt, num_stores=until_turn_turn_by_turn - start_tbt,
                    skip=1, out_address=0, min_particle_id=min_particle_id,
                    max_particle_id=max_particle_id, is_rolling=False,
                    is_turn_ordered=True, cbuffer=beam_elements_buffer)
        added += 1
    start_out = max(start_tbt, until_turn_turn_by_turn)
    if until_turn > start_out:
        if skip_turns <= 0:
            skip_turns = 1
        num_stores = until_turn - start_out
        rem = num_stores % skip_turns
        num_stores //= skip_turns
        if rem > 0:
            num_stores += 1
        BeamMonitor(start=start_out, num_stores=num_stores, skip=skip_turns,
                    out_address=0, min_particle_id=min_particle_id,
                    max_particle_id=max_particle_id, is_rolling=True,
                    is_turn_ordered=True, cbuffer=beam_elements_buffer)
        added += 1
    return added


class LimitRect(CObject):
    """be_limit/be_limit_rect.h:26-33; beam_elements.py:370-428"""
    _typeid = 11
    min_x = CField(0, "float64", default=-1.0)
    max_x = CField(1, "float64", default=+1.0)
    min_y = CField(2, "float64", default=-1.0)
    max_y = CField(3, "float64", default=+1.0)

    def __init__(self, min_x=None, max_x=None, min_y=None, max_y=None,
                 min_coord=-1.0, max_coord=1.0, **kwargs):
        min_x = min_coord if min_x is None else min_x
        min_y = min_coord if min_y is None else min_y
        max_x = max_coord if max_x is None else max_x
        max_y = max_coord if max_y is None else max_y
        if min_x > max_x:
            min_x, max_x = max_x, min_x
        if min_y > max_y:
            min_y, max_y = max_y, min_y
        if (max_x - min_x) * (max_y - min_y) <= 0.0:
            raise ValueError("min_x, max_x, min_y, max_y have to delimit a "
                             "non-vanishing rectangle")
        super().__init__(min_x=min_x, max_x=max_x, min_y=min_y, max_y=max_y,
                         **kwargs)


class LimitEllipse(CObject):
    """be_limit/be_limit_ellipse.h:28-33; beam_elements.py:431-467"""
    _typeid = 12
    a_squ = CField(0, "float64", default=+1.0)
    b_squ = CField(1, "float64", default=+1.0)
    a_b_squ = CField(2, "float64", default=+1.0)

    def __init__(self, a_squ=None, b_squ=None, **kwargs):
        a = kwargs.pop("a", None)
        b = kwargs.pop("b", None)
        if a_squ is None and a is not None and a > 0.0:
            a_squ = a * a
        if b_squ is None and b is not None and b > 0.0:
            b_squ = b * b
        a_squ = 1.0 if a_squ is None else a_squ
        b_squ = 1.0 if b_squ is None else b_squ
        if not (a_squ > 0.0 and b_squ > 0.0):
            raise ValueError("a_squ and b_squ have to be positive definite")
        super().__init__(a_squ=a_squ, b_squ=b_squ, a_b_squ=a_squ * b_squ,
                         **kwargs)

    def set_half_axes(self, a, b):
        return self.set_half_axes_squ(a * a, b * b)

    def set_half_axes_squ(self, a_squ, b_squ):
        self.a_squ = a_squ
        self.b_squ = b_squ
        self.a_b_squ = a_squ * b_squ
        return self


class LimitRectEllipse(CObject):
    """be_limit/be_limit_rect_ellipse.h:26-33; beam_elements.py:600-654"""
    _typeid = 16
    max_x = CField(0, "float64", default=+1.0)
    max_y = CField(1, "float64", default=+1.0)
    a_squ = CField(2, "float64", default=+1.0)
    b_squ = CField(3, "float64", default=+1.0)
    a_b_squ = CField(4, "float64", default=+1.0)

    def __init__(self, max_x=None, max_y=None, a_squ=None, b_squ=None,
                 **kwargs):
        a = kwargs.pop("a", None)
        b = kwargs.pop("b", None)
        max_x = 1.0 if max_x is None else max_x
        max_y = 1.0 if max_y is None else max_y
        if a_squ is None and a is not None and a > 0.0:
            a_squ = a * a
        if b_squ is None and b is not None and b > 0.0:
            b_squ = b * b
        a_squ = 1.0 if a_squ is None else a_squ
        b_squ = 1.0 if b_squ is None else b_squ
        if max_x < 0.0 or max_y < 0.0:
            raise ValueError("max_x / max_y have to be positive definite")
        if a_squ < 0.0 or b_squ < 0.0:
            raise ValueError("a_squ and b_squ have to be positive definite")
        super().__init__(max_x=max_x, max_y=max_y, a_squ=a_squ, b_squ=b_squ,
                         a_b_squ=a_squ * b_squ, **kwargs)


class DipoleEdge(CObject):
    """be_dipedge/be_dipedge.h:29-33; beam_elements.py:657-702"""
    _typeid = 24
    r21 = CField(0, "float64", default=0.0)
    r43 = CField(1, "float64", default=0.0)

    def __init__(self, r21=None, r43=None, h=None, e1=None, hgap=None,
                 fint=None, **kwargs):
        if r21 is None and r43 is None:
            h = 0.0 if h is None else h
            e1 = 0.0 if e1 is None else e1
            hgap = 0.0 if hgap is None else hgap
            fint = 0.0 if fint is None else fint
            if np.isclose(abs(np.cos(e1)), 0.0):
                raise ValueError("e1 too close to (2k+1) pi/2")
            corr = 2.0 * h * hgap * fint
            r21 = h * np.tan(e1)
            temp = corr / np.cos(e1) * (1.0 + np.sin(e1) * np.sin(e1))
            r43 = -h * np.tan(e1 - temp)
        if r21 is None or r43 is None:
            raise ValueError("DipoleEdge needs r21 and r43 or h, e1, hgap, fint")
        super().__init__(r21=r21, r43=r43, **kwargs)


class BeamBeam4D(CObject):
    """be_beamfields/be_beamfields.h:36-56; beam_elements.py:470-493.
    data = [q_part, N_part, sigma_x, sigma_y, beta_s, min_sigma_diff,
            Delta_x, Delta_y, Dpx_sub, Dpy_sub, enabled]"""
    _typeid = 8
    size = CField(0, "uint64", const=True, default=0)
    data = CField(1, "float64", default=0.0, length="size", pointer=True)

    def __init__(self, **kwargs):
        if "x_bb" in kwargs:
            slots = ("charge", "sigma_x", "sigma_y", "beta_r",
                     "min_sigma_diff", "x_bb", "y_bb", "d_px", "d_py",
                     "enabled")
            data = [QELEM_SCIPY] + [kwargs.pop(s) for s in slots]
            super().__init__(size=len(data), data=data, **kwargs)
        else:
            super().__init__(**kwargs)


def bb6d_data(q_part, phi, alpha, delta_x, delta_y, N_part_per_slice,
              z_slices, Sig_11_0, Sig_12_0, Sig_13_0, Sig_14_0, Sig_22_0,
              Sig_23_0, Sig_24_0, Sig_33_0, Sig_34_0, Sig_44_0, x_CO=0.0,
              px_CO=0.0, y_CO=0.0, py_CO=0.0, sigma_CO=0.0, delta_CO=0.0,
              min_sigma_diff=1e-10, threshold_singular=1e-28, Dx_sub=0.0,
              Dpx_sub=0.0, Dy_sub=0.0, Dpy_sub=0.0, Dsigma_sub=0.0,
              Ddelta_sub=0.0, enabled=1):
    """Word image of NS(BB6D_data) (be_beamfields.h:202-250) for a
    Gaussian strong beam cut into slices.

    The reference delegates this to the third-party pysixtrack
    ``BB6Ddata.BB6D_init`` (python/sixtracklib/beam_elements.py:501-543),
    which is absent; this is a restatement of the published method
    (Hirata's boosted frame: boost the strong-beam sigma matrix by the
    crossing angle, order slices head first, boost each slice centroid).
    Given the SixTrack inputs of the reference's three BeamBeam6D lenses
    (tests/testdata/bbsimple and beambeam, fort.3 and beam_elements.bin),
    it reproduces pysixtrack's words bit for bit; the closed orbit and the
    subtracted kick are inputs here, as pysixtrack computes them from the
    machine.
    """
    sphi, cphi, tphi = math.sin(phi), math.cos(phi), math.tan(phi)
    salpha, calpha = math.sin(alpha), math.cos(alpha)
    sig = [Sig_11_0, Sig_12_0 / cphi, Sig_13_0, Sig_14_0 / cphi,
           Sig_22_0 / cphi / cphi, Sig_23_0 / cphi, Sig_24_0 / cphi / cphi,
           Sig_33_0, Sig_34_0 / cphi, Sig_44_0 / cphi / cphi]
    z = np.asarray(z_slices, dtype=float)
    npart = np.asarray(N_part_per_slice, dtype=float)
    order = np.argsort(z)[::-1]
    z, npart = z[order], npart[order]
    # boost of (0, 0, 0, 0, sigma=z, 0): h = 0 so the boost is linear
    x_st = calpha * tphi * z
    y_st = salpha * tphi * z
    s_st = z / cphi
    n = len(z)
    head = [q_part, sphi, cphi, tphi, salpha, calpha] + sig + [
        min_sigma_diff, threshold_singular, None, delta_x, delta_y, x_CO,
        px_CO, y_CO, py_CO, sigma_CO, delta_CO, Dx_sub, Dpx_sub, Dy_sub,
        Dpy_sub, Dsigma_sub, Ddelta_sub, ("int", int(enabled))]
    words = np.zeros(38 + 4 * n, dtype=np.uint64)
    fv = words.view(np.float64)
    for k, v in enumerate(head):
        if v is None:
            words[k] = n                     # N_slices (int64)
        elif isinstance(v, tuple):
            words[k] = v[1]                  # enabled flag stored as integer
        else:
            fv[k] = v
    # relative word offsets of the 4 slice arrays (SIXTRL_BB_GET_PTR,
    # be_beamfields.h:23-27): array = &field + field + 1
    for j in range(4):
        words[34 + j] = (38 + j * n) - (34 + j) - 1
    fv[38: 38 + n] = npart
    fv[38 + n: 38 + 2 * n] = x_st
    fv[38 + 2 * n: 38 + 3 * n] = y_st
    fv[38 + 3 * n: 38 + 4 * n] = s_st
    return words


class BeamBeam6D(CObject):
    """be_beamfields/be_beamfields.h:194-250; beam_elements.py:496-543"""
    _typeid = 9
    size = CField(0, "uint64", const=True, default=0)
    data = CField(1, "float64", default=0.0, length="size", pointer=True)

    def __init__(self, **kwargs):
        if "x_bb_co" in kwargs:
            p = kwargs
            words = bb6d_data(
                q_part=QELEM_SCIPY, phi=p.pop("phi"), alpha=p.pop("alpha"),
                delta_x=p.pop("x_bb_co"), delta_y=p.pop("y_bb_co"),
                N_part_per_slice=p.pop("charge_slices"),
                z_slices=p.pop("zeta_slices"),
                Sig_11_0=p.pop("sigma_11"), Sig_12_0=p.pop("sigma_12"),
                Sig_13_0=p.pop("sigma_13"), Sig_14_0=p.pop("sigma_14"),
                Sig_22_0=p.pop("sigma_22"), Sig_23_0=p.pop("sigma_23"),
                Sig_24_0=p.pop("sigma_24"), Sig_33_0=p.pop("sigma_33"),
                Sig_34_0=p.pop("sigma_34"), Sig_44_0=p.pop("sigma_44"),
                x_CO=p.pop("x_co"), px_CO=p.pop("px_co"), y_CO=p.pop("y_co"),
                py_CO=p.pop("py_co"), sigma_CO=p.pop("zeta_co"),
                delta_CO=p.pop("delta_co"),
                min_sigma_diff=p.pop("min_sigma_diff"),
                threshold_singular=p.pop("threshold_singular"),
                Dx_sub=p.pop("d_x"), Dpx_sub=p.pop("d_px"),
                Dy_sub=p.pop("d_y"), Dpy_sub=p.pop("d_py"),
                Dsigma_sub=p.pop("d_zeta"), Ddelta_sub=p.pop("d_delta"),
                enabled=p.pop("enabled"))
            super().__init__(size=len(words), data=words.view(np.float64),
                             **kwargs)
        else:
            super().__init__(**kwargs)


class SCCoasting(CObject):
    """be_beamfields/be_beamfields.h:418-429"""
    _typeid = 13
    number_of_particles = CField(0, "float64", default=0.0)
    circumference = CField(1, "float64", default=1.0)
    sigma_x = CField(2, "float64", default=1.0)
    sigma_y = CField(3, "float64", default=1.0)
    length = CField(4, "float64", default=0.0)
    x_co = CField(5, "float64", default=0.0)
    y_co = CField(6, "float64", default=0.0)
    min_sigma_diff = CField(7, "float64", default=1e-10)
    enabled = CField(8, "uint64", default=1)


def q_gauss_cq(q):
    """internal/math_qgauss.h:137-159 (host-side normalisation constant)."""
    eps = 1e-6
    cq = math.sqrt(math.pi)
    if q >= 1.0 + eps:
        qm1 = q - 1.0
        cq *= math.gamma((3.0 - q) / (2.0 * qm1))
        cq /= math.sqrt(qm1) * math.gamma(1.0 / qm1)
    elif q <= 1.0 - eps:
        omq = 1.0 - q
        tmq = 3.0 - q
        cq *= 2.0 * math.gamma(1.0 / omq)
        cq /= tmq * math.sqrt(omq) * math.gamma(tmq / (2.0 * omq))
    return cq


class SCQGaussProfile(CObject):
    """be_beamfields/be_beamfields.h:638-651"""
    _typeid = 14
    number_of_particles = CField(0, "float64", default=0.0)
    bunchlength_rms = CField(1, "float64", default=1.0)
    sigma_x = CField(2, "float64", default=1.0)
    sigma_y = CField(3, "float64", default=1.0)
    length = CField(4, "float64", default=0.0)
    x_co = CField(5, "float64", default=0.0)
    y_co = CField(6, "float64", default=0.0)
    min_sigma_diff = CField(7, "float64", default=1e-10)
    q_param = CField(8, "float64", default=1.0)
    cq = CField(9, "float64", default=math.sqrt(math.pi))
    enabled = CField(10, "uint64", default=1)

    def __init__(self, **kwargs):
        q = kwargs.get("q_param", 1.0)
        kwargs["cq"] = q_gauss_cq(q)
        kwargs["q_param"] = q
        super().__init__(**kwargs)


class SCInterpolatedProfile(CObject):
    """be_beamfields/be_beamfields.h:1281-1293"""
    _typeid = 257
    number_of_particles = CField(0, "float64", default=0.0)
    sigma_x = CField(1, "float64", default=1.0)
    sigma_y = CField(2, "float64", default=1.0)
    length = CField(3, "float64", default=0.0)
    x_co = CField(4, "float64", default=0.0)
    y_co = CField(5, "float64", default=0.0)
    interpol_data_addr = CField(6, "uint64", default=0)
    line_density_prof_fallback = CField(7, "float64", default=1.0)
    min_sigma_diff = CField(8, "float64", default=1e-10)
    enabled = CField(9, "uint64", default=1)


class LineDensityProfileData(CObject):
    """be_beamfields/be_beamfields.h:887-896; python/sixtracklib/
    sc_line_density.py:31-112.  Line density on the grid z0 + dz * i,
    i < num_values: ``values`` and ``derivatives`` arrays of ``capacity``;
    ``method`` 0 linear (derivatives = forward differences), 1 cubic
    (derivatives = second derivatives of the natural cubic spline), 255
    none (the element falls back to its constant density)."""
    _typeid = 258
    INTERPOL_LINEAR = 0
    INTERPOL_CUBIC = 1
    INTERPOL_NONE = 255

    method = CField(0, "int64", default=0)
    capacity = CField(1, "int64", default=2, const=True)
    values = CField(2, "real", default=0.0, pointer=True, length="capacity")
    derivatives = CField(3, "real", default=0.0, pointer=True,
                         length="capacity")
    z0 = CField(4, "real", default=0.0)
    dz = CField(5, "real", default=1.0)
    num_values = CField(6, "int64", default=0)

    def __init__(self, method=None, capacity=None, num_values=None, **kwargs):
        if isinstance(method, str):
            method = {"linear": self.INTERPOL_LINEAR, "cubic": self.INTERPOL_CUBIC}.get(
                method.lower().strip(), self.INTERPOL_NONE)
        if method is not None:
            kwargs["method"] = method
        if capacity is not None or num_values is not None:
            capacity = 1 if capacity is None else capacity
            num_values = capacity if num_values is None else num_values
            kwargs["capacity"] = max(capacity, num_values)
            kwargs["num_values"] = num_values
        super().__init__(**kwargs)

    @staticmethod
    def ptr_offset():
        """Byte offset an address assignment points at (the object itself)."""
        return 0

    def prepare_interpolation(self):
        """Fill ``derivatives`` for the object's method
        (internal/math_interpol.h:322-600, default boundaries): forward
        differences for linear; for cubic the second derivatives of the
        natural spline, from the tridiagonal system solved by forward
        elimination and back substitution in that file's order."""
        n = int(self.num_values)
        y = np.array(self.values[:n], dtype=np.float64)
        dx = float(self.dz)
        if not (n > 1 and dx > 0.0 and math.isfinite(dx) and math.isfinite(float(self.z0))):
            raise RuntimeError("unable to calculate derivatives for interpolation")
        out = np.array(self.derivatives, dtype=np.float64)
        if int(self.method) == self.INTERPOL_LINEAR:
            out[: n - 1] = (y[1:] - y[:-1]) / dx
        elif int(self.method) == self.INTERPOL_CUBIC:
            if n <= 3:
                raise RuntimeError("cubic interpolation needs more than 3 values")
            out[:n] = _natural_spline_ypp(y, dx)
        else:
            raise RuntimeError("unable to calculate derivatives for interpolation")
        self.derivatives = out

    def interpol(self, z):
        """Interpolated line density at ``z`` (be_beamfields.h:3385-3410):
        0 outside [z0, z0 + dz (num_values - 1))."""
        n, z0, dz = int(self.num_values), float(self.z0), float(self.dz)
        if n <= 0 or int(self.method) not in (self.INTERPOL_LINEAR, self.INTERPOL_CUBIC):
            return 0.0
        if not (z0 <= z <= z0 + dz * n):
            return 0.0
        i = int((z - z0) / dz)
        if i + 1 >= n:
            return 0.0
        y, yd = self.values, self.derivatives
        if int(self.method) == self.INTERPOL_LINEAR:
            return float(y[i] + (z - (z0 + dz * i)) * yd[i])
        t2 = (z0 + dz * (i + 1) - z) / dz
        t1 = (z - (z0 + dz * i)) / dz
        r = y[i + 1] * t1 + y[i] * t2
        r -= yd[i] * dz * dz * (t2 - t2 * t2 * t2) / 6.0
        r -= yd[i + 1] * dz * dz * (t1 - t1 * t1 * t1) / 6.0
        return float(r)


def _natural_spline_ypp(y, dx):
    """Second derivatives at the nodes of the natural cubic spline through
    y on a grid of step dx (internal/math_interpol.h:369-600: h = dx,
    g = 4 dx, right-hand side 6 (forward - backward difference quotient),
    Thomas elimination, then zero second derivative at both ends)."""
    n = len(y)
    g = [0.0] * n
    dy = [0.0] * n
    for i in range(1, n - 1):
        g[i] = (dx + dx) * 2.0
        dy[i] = ((y[i + 1] - y[i]) / dx - (y[i] - y[i - 1]) / dx) * 6.0
    d = [0.0] * n
    ypp = [0.0] * n
    d[0] = g[1]
    ypp[0] = dy[1]
    for i in range(1, n - 2):
        l_i = dx / d[i - 1]
        d[i] = g[i + 1] - l_i * dx
        ypp[i] = dy[i + 1] - ypp[i - 1] * l_i
    i = n - 3
    ypp[i + 1] = ypp[i] / d[i]
    for i in range(n - 4, -1, -1):
        ypp[i + 1] = (ypp[i] - dx * ypp[i + 2]) / d[i]
    ypp[0] = 0.0
    ypp[n - 1] = 0.0
    return np.array(ypp)


class Elements:
    """Lattice container: python/sixtracklib/beam_elements.py:705-790."""

    element_types = {
        "Drift": Drift, "DriftExact": DriftExact, "Multipole": Multipole,
        "RFMultipole": RFMultipole, "Cavity": Cavity, "XYShift": XYShift,
        "SRotation": SRotation, "BeamMonitor": BeamMonitor,
        "LimitRect": LimitRect, "LimitEllipse": LimitEllipse,
        "LimitRectEllipse": LimitRectEllipse, "DipoleEdge": DipoleEdge,
        "BeamBeam4D": BeamBeam4D, "BeamBeam6D": BeamBeam6D,
        "SCCoasting": SCCoasting, "SCQGaussProfile": SCQGaussProfile,
        "SCInterpolatedProfile": SCInterpolatedProfile,
    }

    def __init__(self, cbuffer=None):
        self.cbuffer = cbuffer if cbuffer is not None else CBuffer()
        for name, cls in self.element_types.items():
            setattr(self, name, self._mk_fun(cls))

    def _mk_fun(self, cls):
        def fun(*args, **kwargs):
            return cls(*args, cbuffer=self.cbuffer, **kwargs)
        return fun

    @classmethod
    def fromfile(cls, filename):
        return cls(cbuffer=CBuffer.fromfile(filename))

    def tofile(self, filename):
        self.cbuffer.tofile(filename)

    to_file = tofile

    def get_elements(self):
        return [get_object(self.cbuffer, i)
                for i in range(self.cbuffer.n_objects)]

    def get(self, objid):
        return get_object(self.cbuffer, objid)

    def __len__(self):
        return self.cbuffer.n_objects

"""Particle sets: the SoA ``NS(Particles)`` object in a CObjects buffer.

Mirror of python/sixtracklib/particles.py (``Particles``, ``ParticlesSet``,
``makeCopy``, ``calcParticlesDifference``, ``compareParticlesDifference``)
on top of this package's own CBuffer.  Layout: count + 21 field pointers,
17 FP64 fields then 4 int64 fields, each field a contiguous array
(sixtracklib/common/particles.h:38-67) -- already the coalescing-friendly
structure-of-arrays the GPU kernels read.
"""
from __future__ import annotations

import numpy as np

from .cobject import CObject, CField, get_object
from .cobjects import CBuffer

REAL_FIELDS = ("q0", "mass0", "beta0", "gamma0", "p0c", "s", "x", "y", "px",
               "py", "zeta", "psigma", "delta", "rpp", "rvv", "chi",
               "charge_ratio")
INDEX_FIELDS = ("particle_id", "at_element", "at_turn", "state")
ALL_FIELDS = REAL_FIELDS + INDEX_FIELDS

PROTON_MASS_EV = 938272088.16   # internal/physics_constants.h:53


def update_delta_arrays(delta, beta0):
    """rvv, rpp, psigma for a new delta exactly as the C tracker derives
    them (sixtracklib/common/particles.h:5459-5497), evaluated with IEEE
    float64 numpy operations in the same order."""
    delta = np.asarray(delta, dtype=np.float64)
    beta0 = np.asarray(beta0, dtype=np.float64)
    delta_beta0 = delta * beta0
    ptau_beta0 = np.sqrt(delta_beta0 * delta_beta0 +
                         2.0 * delta_beta0 * beta0 + 1.0) - 1.0
    one_plus_delta = 1.0 + delta
    rvv = one_plus_delta / (1.0 + ptau_beta0)
    rpp = 1.0 / one_plus_delta
    psigma = ptau_beta0 / (beta0 * beta0)
    return rvv, rpp, psigma


class Particles(CObject):
    """python/sixtracklib/particles.py:6-225"""

    pmass = 938.2720813e6

    def _set_p0c(self):
        energy0 = np.sqrt(self.p0c ** 2 + self.mass0 ** 2)
        self.beta0 = self.p0c / energy0
        self.gamma0 = energy0 / self.mass0

    def _set_delta(self):
        rep = np.sqrt(self.delta ** 2 + 2 * self.delta + 1 / self.beta0 ** 2)
        irpp = 1 + self.delta
        self.rpp = 1 / irpp
        beta = irpp / rep
        self.rvv = beta / self.beta0
        self.psigma = (np.sqrt(self.delta ** 2 + 2 * self.delta +
                               1 / self.beta0 ** 2) / self.beta0 -
                       1 / self.beta0 ** 2)

    _typeid = 1
    num_particles = CField(0, "int64", const=True)
    q0 = CField(1, "real", length="num_particles", default=0.0, pointer=True)
    mass0 = CField(2, "real", length="num_particles", default=pmass, pointer=True)
    beta0 = CField(3, "real", length="num_particles", default=1.0, pointer=True)
    gamma0 = CField(4, "real", length="num_particles", default=1.0, pointer=True)
    p0c = CField(5, "real", length="num_particles", default=1.0, pointer=True,
                 setter=_set_p0c)
    s = CField(6, "real", length="num_particles", default=0.0, pointer=True)
    x = CField(7, "real", length="num_particles", default=0.0, pointer=True)
    y = CField(8, "real", length="num_particles", default=0.0, pointer=True)
    px = CField(9, "real", length="num_particles", default=0.0, pointer=True)
    py = CField(10, "real", length="num_particles", default=0.0, pointer=True)
    zeta = CField(11, "real", length="num_particles", default=0.0, pointer=True)
    psigma = CField(12, "real", length="num_particles", default=0.0, pointer=True)
    delta = CField(13, "real", length="num_particles", default=0.0, pointer=True,
                   setter=_set_delta)
    rpp = CField(14, "real", length="num_particles", default=1.0, pointer=True)
    rvv = CField(15, "real", length="num_particles", default=1.0, pointer=True)
    chi = CField(16, "real", length="num_particles", default=1.0, pointer=True)
    charge_ratio = CField(17, "real", length="num_particles", default=1.0,
                          pointer=True)
    particle_id = CField(18, "int64", length="num_particles", default=0,
                         pointer=True)
    at_element = CField(19, "int64", length="num_particles", default=0,
                        pointer=True)
    at_turn = CField(20, "int64", length="num_particles", default=0,
                     pointer=True)
    state = CField(21, "int64", length="num_particles", default=1,
                   pointer=True)

    @classmethod
    def from_ref(cls, num_particles=1, mass0=pmass, p0c=1e9, q0=1,
                 cbuffer=None):
        return cls(num_particles=num_particles,
                   particle_id=np.arange(num_particles),
                   cbuffer=cbuffer).set_reference(mass0=mass0, p0c=p0c, q0=q0)

    sigma = property(lambda self: (self.beta0 / self.beta) * self.zeta)
    beta = property(lambda p: (1 + p.delta) / (1 / p.beta0 + p.ptau))

    @property
    def ptau(self):
        return (np.sqrt(self.delta ** 2 + 2 * self.delta + 1 / self.beta0 ** 2)
                - 1 / self.beta0)

    def set_reference(self, p0c=7e12, mass0=pmass, q0=1):
        self.q0 = q0
        self.mass0 = mass0
        self.p0c = p0c
        return self

    def update_delta(self, delta):
        """Set delta and derive rvv/rpp/psigma with the C tracker's formula
        (particles.h:5459-5497), bit-identical to it."""
        d = np.broadcast_to(np.asarray(delta, dtype=np.float64),
                            (self.num_particles,)).copy()
        rvv, rpp, psigma = update_delta_arrays(d, np.array(self.beta0))
        # write through the view: the ``delta`` setter would re-derive the
        # energy variables with the Python formula instead
        type(self).delta._view(self)[:] = d
        self.rvv = rvv
        self.rpp = rpp
        self.psigma = psigma
        return self

    def field(self, name):
        if name == "at_element_id":
            name = "at_element"
        return getattr(self, name)

    def to_dict(self):
        return {name: np.array(getattr(self, name)) for name in ALL_FIELDS}

    def copy_from(self, other, *, src_index=None):
        for name in ALL_FIELDS:
            getattr(self, name)[:] = getattr(other, name)
        return self


def makeCopy(orig, cbuffer=None):
    p = Particles(cbuffer=cbuffer, num_particles=orig.num_particles)
    return p.copy_from(orig)


def calcParticlesDifference(lhs, rhs, cbuffer=None):
    diff = Particles(cbuffer=cbuffer, num_particles=lhs.num_particles)
    for name in ALL_FIELDS:
        getattr(diff, name)[:] = getattr(lhs, name) - getattr(rhs, name)
    return diff


def compareParticlesDifference(lhs, rhs, abs_treshold=None):
    """0 if equal (or within ``abs_treshold`` on the real fields and exact on
    the index fields), following tests/sixtracklib/testlib/common/particles/
    particles.h:760-930."""
    if lhs.num_particles != rhs.num_particles:
        return -1
    for name in REAL_FIELDS:
        a, b = getattr(lhs, name), getattr(rhs, name)
        if abs_treshold is None:
            if not np.array_equal(a, b):
                return -1
        elif np.any(np.abs(a - b) > abs_treshold):
            return -1
    for name in INDEX_FIELDS:
        if not np.array_equal(getattr(lhs, name), getattr(rhs, name)):
            return -1
    return 0


class ParticlesSet:
    """A CBuffer holding several Particles objects
    (python/sixtracklib/particles.py ``ParticlesSet``): ``particles`` lists
    the objects of the buffer, so a set made on an existing buffer (a job's
    ``output``, a file) sees every Particles object in it."""

    element_types = {"Particles": Particles}

    def __init__(self, cbuffer=None):
        self.cbuffer = cbuffer if cbuffer is not None else CBuffer()

    @property
    def particles(self):
        cb = self.cbuffer
        return [Particles._from_buffer(cb, i) for i in range(cb.n_objects)
                if cb.get_object_typeid(i) == Particles._typeid]

    def Particles(self, **kwargs):
        return Particles(cbuffer=self.cbuffer, **kwargs)

    @classmethod
    def fromfile(cls, filename):
        return cls(cbuffer=CBuffer.fromfile(filename))

    def tofile(self, filename):
        self.cbuffer.tofile(filename)

    to_file = tofile

    @property
    def num_particles(self):
        return sum(p.num_particles for p in self.particles)


class ParticlesAddr(CObject):
    """NS(ParticlesAddr) (particles/particles_addr.h:27-55; Python
    python/sixtracklib/particles.py ``ParticlesAddr``): the particle count
    and the device address of each of the 21 field arrays of one particle
    set, object type 512.  A job fills such records with
    ``fetch_particle_addresses``."""
    _typeid = 512
    num_particles = CField(0, "int64", default=0)
    q0 = CField(1, "uint64", default=0)
    mass0 = CField(2, "uint64", default=0)
    beta0 = CField(3, "uint64", default=0)
    gamma0 = CField(4, "uint64", default=0)
    p0c = CField(5, "uint64", default=0)
    s = CField(6, "uint64", default=0)
    x = CField(7, "uint64", default=0)
    y = CField(8, "uint64", default=0)
    px = CField(9, "uint64", default=0)
    py = CField(10, "uint64", default=0)
    zeta = CField(11, "uint64", default=0)
    psigma = CField(12, "uint64", default=0)
    delta = CField(13, "uint64", default=0)
    rpp = CField(14, "uint64", default=0)
    rvv = CField(15, "uint64", default=0)
    chi = CField(16, "uint64", default=0)
    charge_ratio = CField(17, "uint64", default=0)
    particle_id = CField(18, "uint64", default=0)
    at_element = CField(19, "uint64", default=0)
    at_turn = CField(20, "uint64", default=0)
    state = CField(21, "uint64", default=0)


def particles_in(cbuffer, index=0):
    """The ``index``-th object of ``cbuffer`` (a CBuffer, or a ParticlesSet
    such as a job's ``output``) as :class:`Particles`."""
    if isinstance(cbuffer, ParticlesSet):
        cbuffer = cbuffer.cbuffer
    if cbuffer.get_object_typeid(index) != Particles._typeid:
        raise TypeError("object %d is not a Particles object" % index)
    return Particles._from_buffer(cbuffer, index)

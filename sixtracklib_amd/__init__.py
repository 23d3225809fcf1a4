"""sixtracklib_amd -- MI355X-native particle tracking with the SixTrackLib API.

Pure-Python parts (buffers, particles, elements, lattice generators) import
without the native library; ``TrackJob`` and the GPU front ends load
libsixtrack_amd.so on first use and raise if it is missing.
"""
from .cobjects import CBuffer
from .cobject import CObject, CField, get_object
from .particles import (Particles, ParticlesSet, ParticlesAddr, makeCopy,
                        calcParticlesDifference, compareParticlesDifference)
from .beam_elements import (Drift, DriftExact, Multipole, RFMultipole, Cavity, XYShift,
                            SRotation, BeamMonitor, LimitRect, LimitEllipse,
                            LimitRectEllipse, DipoleEdge, BeamBeam4D, BeamBeam6D,
                            SCCoasting, SCQGaussProfile, SCInterpolatedProfile,
                            LineDensityProfileData, Elements,
                            append_beam_monitors_to_lattice)

__all__ = ["CBuffer", "CObject", "CField", "get_object", "Particles", "ParticlesSet", "ParticlesAddr",
           "makeCopy", "calcParticlesDifference", "compareParticlesDifference",
           "Drift", "DriftExact", "Multipole", "RFMultipole", "Cavity", "XYShift",
           "SRotation", "BeamMonitor", "LimitRect", "LimitEllipse", "LimitRectEllipse",
           "DipoleEdge", "BeamBeam4D", "BeamBeam6D", "SCCoasting", "SCQGaussProfile",
           "SCInterpolatedProfile", "LineDensityProfileData", "Elements",
           "append_beam_monitors_to_lattice", "TrackJob", "TrackJobBaseNew", "CudaTrackJob",
           "TrackBatch", "Buffer", "AssignAddressItem", "supports"]


def __getattr__(name):
    if name == "TrackJob":
        from .trackjob import TrackJob
        return TrackJob
    if name == "TrackJobBaseNew":
        from .trackjob import TrackJobBaseNew
        return TrackJobBaseNew
    if name == "CudaTrackJob":
        from .cuda import CudaTrackJob
        return CudaTrackJob
    if name in ("Buffer", "AssignAddressItem"):
        from . import buffer
        return getattr(buffer, name)
    if name == "supports":
        from .config_helper import supports
        return supports
    if name == "TrackBatch":
        from .trackbatch import TrackBatch
        return TrackBatch
    raise AttributeError(name)

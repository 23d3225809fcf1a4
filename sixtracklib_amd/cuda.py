"""The CUDA track job of the reference's Python binding
(python/sixtracklib/cuda.py ``CudaTrackJob``), running the HIP job.

A script written for the reference's GPU job constructs
``CudaTrackJob(lattice, particles, ...)``: here that is the same
``TrackJobBaseNew`` on an AMD GPU.  The node id string selects the HIP
device ("0.N" or "N").  The reference's controller and argument objects
(cuda/controller.h, cuda/argument.h) have no counterpart: ``has_controller``
and the ``has_*_arg`` properties are False, and ``controller`` raises.
"""
from __future__ import annotations

from . import stcommon as st
from .trackjob import TrackJobBaseNew

__all__ = ["CudaTrackJob"]


def _node_config(node_id_str, config_str):
    cfg = (config_str or "").strip()
    if node_id_str:
        dev = str(node_id_str).split(".")[-1].strip()
        cfg = (cfg + " device=%d" % int(dev)).strip()
    return cfg


class CudaTrackJob(TrackJobBaseNew):
    def __init__(self, beam_elements_buffer, particles_buffer, particle_set_index=0,
                 until_turn_elem_by_elem=0, output_buffer=None, node_id_str=None,
                 config_str=None):
        cfg = _node_config(node_id_str, config_str)
        ptr = (st.st_CudaTrackJob_new_from_config_str(cfg.encode()) if cfg
               else st.st_CudaTrackJob_create())
        if not ptr:
            raise RuntimeError("Error while creating CudaTrackJob: " + st.last_error())
        super().__init__(ptr_track_job=ptr, owns_ptr=True)
        self._config_str = cfg
        self._reset_detailed(beam_elements_buffer, particles_buffer, particle_set_index,
                             until_turn_elem_by_elem, output_buffer)

    @property
    def has_controller(self):
        return bool(st.st_CudaTrackJob_has_controller(self.ptr_st_track_job))

    @property
    def controller(self):
        raise RuntimeError("CudaTrackJob on HIP has no controller object")

    @property
    def has_particles_arg(self):
        return bool(st.st_CudaTrackJob_has_particles_arg(self.ptr_st_track_job))

    @property
    def has_beam_elements_arg(self):
        return bool(st.st_CudaTrackJob_has_beam_elements_arg(self.ptr_st_track_job))

    @property
    def has_output_arg(self):
        return bool(st.st_CudaTrackJob_has_output_arg(self.ptr_st_track_job))

    @property
    def has_particles_addr_arg(self):
        return bool(st.st_CudaTrackJob_has_particles_addr_arg(self.ptr_st_track_job))

    @property
    def has_debug_register_arg(self):
        return bool(st.st_CudaTrackJob_has_debug_register_arg(self.ptr_st_track_job))

    @property
    def default_threads_per_block(self):
        return int(st.st_CudaTrackJob_default_threads_per_block(self.ptr_st_track_job))

    @property
    def default_track_threads_per_block(self):
        return int(st.st_CudaTrackJob_default_track_threads_per_block(self.ptr_st_track_job))

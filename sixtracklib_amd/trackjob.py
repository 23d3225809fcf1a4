"""Python ``TrackJob``: drop-in for ``sixtracklib.TrackJob``
(python/sixtracklib/trackjob.py:1213-1560).

Same constructor signature and methods (``track_until``,
``track_elem_by_elem``, ``track_line``, ``collect``, ``collect_particles``,
``collect_output``, ``push_particles``, ``push_beam_elements``, ``output``,
``has_elem_by_elem_output``, ...).  Buffers are this package's
:class:`~sixtracklib_amd.cobjects.CBuffer` (or objects carrying one:
``Elements``, ``ParticlesSet``, ``Particles``); as in the reference, the
job maps the buffers' memory, so ``collect()`` updates the Python-side
arrays in place.  Tracking runs on the GPU only: ``arch`` is "hip"; the
reference's "cpu"/"opencl" are rejected with ``ValueError``.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import stcommon as st
from .cobjects import CBuffer
from .particles import Particles, ParticlesSet

_GPU_ARCHS = ("hip", "amd", "amdgpu", "gpu")


def get_cbuffer_from_obj(obj):
    """python/sixtracklib/buffer.py get_cbuffer_from_obj"""
    if isinstance(obj, CBuffer):
        return obj
    for attr in ("cbuffer", "_buffer"):
        cb = getattr(obj, attr, None)
        if isinstance(cb, CBuffer):
            return cb
    raise TypeError("cannot get a CBuffer from %r" % (obj,))


class TrackJob:
    def __init__(self, beam_elements_buffer, particles_buffer,
                 until_turn_elem_by_elem=0, arch="hip", device_id=None,
                 device=None, output_buffer=None, config_str=None,
                 particle_set_indices=None):
        self.ptr_st_track_job = None
        self._c_buffers = []
        if device is not None:
            parts = str(device).split(":", 1)
            arch = parts[0]
            if len(parts) > 1:
                device_id = parts[1]
        arch = str(arch).strip().lower()
        if arch not in _GPU_ARCHS:
            raise ValueError("architecture %r is not available: sixtracklib_amd "
                             "tracks on AMD GPUs only (arch='hip')" % arch)
        cfg = config_str or ""
        if device_id is not None:
            cfg = (cfg + " device=%s" % str(device_id).split(".")[0]).strip()
        self._arch = arch
        self._beam_elements_buffer = get_cbuffer_from_obj(beam_elements_buffer)
        self._particles_buffer = get_cbuffer_from_obj(particles_buffer)
        self._until_turn_ebe = int(until_turn_elem_by_elem)

        self._c_pbuf = self._map(self._particles_buffer)
        self._c_ebuf = self._map(self._beam_elements_buffer)

        if particle_set_indices is None:
            particle_set_indices = [0]
        idx = (st.st_buffer_size_t * len(particle_set_indices))(*particle_set_indices)
        self._set_indices = idx

        self._config_str = cfg
        self._output_buffer, self._c_obuf = self._prepare_output(
            self._c_pbuf, idx, self._c_ebuf, self._until_turn_ebe, output_buffer)

        self.ptr_st_track_job = st.st_TrackJob_new_detailed(
            arch.encode(), self._c_pbuf, len(particle_set_indices), idx,
            self._c_ebuf, self._c_obuf, self._until_turn_ebe,
            st.string_to_encoded_ctypes_str(cfg))
        if not self.ptr_st_track_job:
            self._free_buffers()
            raise RuntimeError("unable to create TrackJob: " + st.last_error())
        self._last_status = 0
        self._last_track_status = 0

    def _prepare_output(self, c_pbuf, idx, c_ebuf, until_turn_ebe, output_buffer):
        """The output buffer (CBuffer, mapped NS(Buffer)*): the caller's, or
        one sized and built on the Python side like the reference binding
        (trackjob.py:1355-1405), or none when nothing is stored."""
        sizes = (ct.c_int64 * 4096)()
        has_ebe = ct.c_int64(0)
        n = st.st_TrackJob_required_output_sets(
            c_pbuf, len(idx), idx, c_ebuf, until_turn_ebe, sizes, 4096, ct.byref(has_ebe),
            st.string_to_encoded_ctypes_str(self._config_str))
        if n < 0:
            raise RuntimeError("cannot determine the output requirements: " + st.last_error())
        ob = None
        if output_buffer is not None:
            ob = get_cbuffer_from_obj(output_buffer)
        elif n > 0:
            total = sum(int(sizes[k]) for k in range(n))
            ob = CBuffer(max_slots=22 * n + 21 * total + 64, max_objects=n + 4,
                         max_pointers=21 * n + 4)
            for k in range(n):
                Particles(cbuffer=ob, num_particles=int(sizes[k]))
        return ob, (self._map(ob) if ob is not None else None)

    def _map(self, cbuf):
        p = st.st_Buffer_new_on_memory(cbuf.pointer, cbuf.words.nbytes)
        if not p:
            raise RuntimeError("cannot map buffer")
        self._c_buffers.append(p)
        return p

    def _free_buffers(self, _buffer_delete=st.st_Buffer_delete):
        for p in self._c_buffers:
            _buffer_delete(p)
        self._c_buffers = []

    # the library functions are bound as defaults: at interpreter exit the
    # module globals may already be gone when __del__ runs
    def __del__(self, _job_delete=st.st_TrackJob_delete):
        if getattr(self, "ptr_st_track_job", None):
            _job_delete(self.ptr_st_track_job)
            self.ptr_st_track_job = None
        if getattr(self, "_c_buffers", None):
            self._free_buffers()

    # -- properties -------------------------------------------------------
    @property
    def last_status(self):
        return self._last_status

    @property
    def last_track_status(self):
        return self._last_track_status

    @property
    def arch(self):
        return self._arch

    arch_str = arch

    @property
    def output_buffer(self):
        return self._output_buffer

    @property
    def output(self):
        """python/sixtracklib/trackjob.py: the output buffer as a
        ParticlesSet (the element-by-element set first, then one set per
        BeamMonitor)."""
        return ParticlesSet(self._output_buffer)

    @property
    def particles_buffer(self):
        return self._particles_buffer

    @property
    def beam_elements_buffer(self):
        return self._beam_elements_buffer

    @property
    def type_str(self):
        return st.st_TrackJob_get_type_str(self.ptr_st_track_job).decode()

    @property
    def num_beam_monitors(self):
        return st.st_TrackJob_get_num_beam_monitors(self.ptr_st_track_job)

    @property
    def has_elem_by_elem_output(self):
        return st.st_TrackJob_has_elem_by_elem_output(self.ptr_st_track_job)

    @property
    def has_beam_monitor_output(self):
        return st.st_TrackJob_has_beam_monitor_output(self.ptr_st_track_job)

    @property
    def elem_by_elem_output_offset(self):
        return st.st_TrackJob_get_elem_by_elem_output_buffer_offset(self.ptr_st_track_job)

    @property
    def beam_monitor_output_offset(self):
        return st.st_TrackJob_get_beam_monitor_output_buffer_offset(self.ptr_st_track_job)

    @property
    def has_output_buffer(self):
        return st.st_TrackJob_has_output_buffer(self.ptr_st_track_job)

    @property
    def requires_collecting(self):
        return True

    @property
    def last_kernel_time_ms(self):
        """device time of the kernels of the last track_* call (HIP events
        on the job's stream)"""
        return st.st_TrackJob_last_kernel_time_ms(self.ptr_st_track_job)

    @property
    def uses_fma(self):
        return bool(st.st_TrackJob_uses_fma(self.ptr_st_track_job))

    @property
    def uses_full_kernel(self):
        """True when the lattice holds maps outside the lean set (monitors,
        beam-beam, space charge, RF multipoles), which the full kernel
        tracks."""
        return bool(st.st_TrackJob_uses_full_kernel(self.ptr_st_track_job))

    # -- tracking ----------------------------------------------------------
    def _check(self, status, what):
        """python/sixtracklib/trackjob.py:1522-1553: the status is kept in
        last_track_status, a failure raises, and the job is returned so
        that calls chain (job.track_until(n).collect())."""
        self._last_track_status = status
        self._last_status = status
        if status != 0:
            raise RuntimeError("%s failed with status %d %s" % (what, status, st.last_error()))
        return self

    def track(self, until_turn):
        return self.track_until(until_turn)

    def track_until(self, until_turn):
        return self._check(st.st_TrackJob_track_until(self.ptr_st_track_job, int(until_turn)),
                           "track_until")

    def track_elem_by_elem(self, until_turn):
        return self._check(st.st_TrackJob_track_elem_by_elem(self.ptr_st_track_job,
                                                             int(until_turn)),
                           "track_elem_by_elem")

    def track_line(self, begin_idx, end_idx, finish_turn=False):
        return self._check(st.st_TrackJob_track_line(self.ptr_st_track_job, int(begin_idx),
                                                     int(end_idx), bool(finish_turn)),
                           "track_line")

    def collect(self):
        st.st_TrackJob_collect(self.ptr_st_track_job)

    def collect_particles(self):
        st.st_TrackJob_collect_particles(self.ptr_st_track_job)

    def collect_beam_elements(self):
        st.st_TrackJob_collect_beam_elements(self.ptr_st_track_job)

    def collect_output(self):
        st.st_TrackJob_collect_output(self.ptr_st_track_job)

    def push_particles(self):
        st.st_TrackJob_push_particles(self.ptr_st_track_job)

    def push_beam_elements(self):
        st.st_TrackJob_push_beam_elements(self.ptr_st_track_job)

    def _push_output(self):
        st.st_TrackJob_push_output(self.ptr_st_track_job)

    # -- stored buffers (python/sixtracklib/trackjob.py:2200-2285) ---------
    @property
    def has_stored_buffers(self):
        return bool(st.st_TrackJob_has_stored_buffers(self.ptr_st_track_job))

    @property
    def num_stored_buffers(self):
        return int(st.st_TrackJob_num_stored_buffers(self.ptr_st_track_job))

    @property
    def min_stored_buffer_id(self):
        return int(st.st_TrackJob_min_stored_buffer_id(self.ptr_st_track_job))

    @property
    def max_stored_buffer_id(self):
        return int(st.st_TrackJob_max_stored_buffer_id(self.ptr_st_track_job))

    def add_stored_buffer(self, cbuffer=None, size=None, ptr_c_buffer=None,
                          take_ownership=False, delete_ptr_after_move=False):
        """Keep ``cbuffer`` (a CBuffer, mapped without a copy), a new empty
        buffer of ``size`` bytes owned by the job, or an existing
        ``NS(Buffer)*``; returns its buffer id (st_ARCH_ILLEGAL_BUFFER_ID
        on failure)."""
        if not hasattr(self, "_stored"):
            self._stored = {}
        bid = st.st_ARCH_ILLEGAL_BUFFER_ID.value
        if cbuffer is not None:
            cb = get_cbuffer_from_obj(cbuffer)
            bid = st.st_TrackJob_add_stored_buffer(self.ptr_st_track_job, self._map(cb),
                                                   False, False)
            if bid != st.st_ARCH_ILLEGAL_BUFFER_ID.value:
                self._stored[bid] = cb
        elif size is not None and size > 0:
            bid = st.st_TrackJob_create_stored_buffer(self.ptr_st_track_job, int(size))
        elif ptr_c_buffer is not None:
            bid = st.st_TrackJob_add_stored_buffer(self.ptr_st_track_job, ptr_c_buffer,
                                                   bool(take_ownership),
                                                   bool(delete_ptr_after_move))
        return int(bid)

    def remove_stored_buffer(self, buffer_id):
        getattr(self, "_stored", {}).pop(buffer_id, None)
        return st.st_TrackJob_remove_stored_buffer(self.ptr_st_track_job, int(buffer_id))

    def owns_stored_buffer(self, buffer_id):
        return bool(st.st_TrackJob_owns_stored_buffer(self.ptr_st_track_job, int(buffer_id)))

    def stored_buffer_is_cbuffer(self, buffer_id):
        return buffer_id in getattr(self, "_stored", {})

    def stored_buffer_cbuffer(self, buffer_id):
        if not self.stored_buffer_is_cbuffer(buffer_id):
            raise RuntimeError("Unable to retrieve CBuffer for buffer_id=%d" % buffer_id)
        return self._stored[buffer_id]

    def ptr_stored_buffer(self, buffer_id):
        return st.st_TrackJob_stored_buffer(self.ptr_st_track_job, int(buffer_id))

    def stored_buffer(self, buffer_id):
        return getattr(self, "_stored", {}).get(buffer_id, None)

    def push_stored_buffer(self, buffer_id):
        self._last_status = st.st_TrackJob_push_stored_buffer(self.ptr_st_track_job,
                                                              int(buffer_id))
        if self._last_status != 0:
            raise RuntimeError("Unable to push stored buffer %d" % buffer_id)
        return self

    def collect_stored_buffer(self, buffer_id):
        self._last_status = st.st_TrackJob_collect_stored_buffer(self.ptr_st_track_job,
                                                                 int(buffer_id))
        if self._last_status != 0:
            raise RuntimeError("Unable to collect stored buffer %d" % buffer_id)
        return self

    # -- address assignments (python/sixtracklib/trackjob.py:1833-2040) ----
    def add_assign_address_item(self, item=None, dest_elem_type_id=None, dest_buffer_id=None,
                                dest_elem_index=None, dest_pointer_offset=None,
                                src_elem_type_id=None, src_buffer_id=None, src_elem_index=None,
                                src_pointer_offset=None):
        ptr = None
        if item is not None:
            if isinstance(item, st.st_AssignAddressItem):
                ptr = st.st_TrackJob_add_assign_address_item(self.ptr_st_track_job,
                                                             ct.byref(item))
            else:                                  # an AssignAddressItem CObject
                it = st.st_AssignAddressItem(*[int(getattr(item, f)) for f, _ in
                                               st.st_AssignAddressItem._fields_])
                ptr = st.st_TrackJob_add_assign_address_item(self.ptr_st_track_job,
                                                             ct.byref(it))
        elif None not in (dest_elem_type_id, dest_buffer_id, dest_elem_index,
                          dest_pointer_offset, src_elem_type_id, src_buffer_id,
                          src_elem_index, src_pointer_offset):
            ptr = st.st_TrackJob_add_assign_address_item_detailed(
                self.ptr_st_track_job, int(dest_elem_type_id), int(dest_buffer_id),
                int(dest_elem_index), int(dest_pointer_offset), int(src_elem_type_id),
                int(src_buffer_id), int(src_elem_index), int(src_pointer_offset))
        if not ptr:
            raise ValueError("unable to add AssignAddressItem given by these parameters")
        return ptr

    def has_assign_items(self, dest_buffer_id, src_buffer_id):
        return bool(st.st_TrackJob_has_assign_items(self.ptr_st_track_job, int(dest_buffer_id),
                                                    int(src_buffer_id)))

    def num_assign_items(self, dest_buffer_id, src_buffer_id):
        return int(st.st_TrackJob_num_assign_items(self.ptr_st_track_job, int(dest_buffer_id),
                                                   int(src_buffer_id)))

    @property
    def total_num_assign_items(self):
        return int(st.st_TrackJob_total_num_assign_items(self.ptr_st_track_job))

    def commit_address_assignments(self):
        self._last_status = st.st_TrackJob_commit_address_assignments(self.ptr_st_track_job)
        if self._last_status != 0:
            raise RuntimeError("Unable to commit address assignment items")
        return self

    def assign_all_addresses(self):
        self._last_status = st.st_TrackJob_assign_all_addresses(self.ptr_st_track_job)
        if self._last_status != 0:
            raise RuntimeError("Unable to perform assignment of all address items: " +
                               st.last_error())
        return self

    def assign_addresses(self, dest_buffer_id, src_buffer_id):
        if dest_buffer_id is None and src_buffer_id is None:
            return self.assign_all_addresses()
        if dest_buffer_id is None or src_buffer_id is None:
            raise ValueError("give both dest_buffer_id and src_buffer_id, or neither")
        self._last_status = st.st_TrackJob_assign_addresses(self.ptr_st_track_job,
                                                            int(dest_buffer_id),
                                                            int(src_buffer_id))
        if self._last_status != 0:
            raise RuntimeError("Unable to perform assignment of address items: " +
                               st.last_error())
        return self

    @staticmethod
    def _item_struct(item):
        if isinstance(item, st.st_AssignAddressItem):
            return item
        return st.st_AssignAddressItem(*[int(getattr(item, f)) for f, _ in
                                         st.st_AssignAddressItem._fields_])

    @staticmethod
    def _detail(*args):
        return None not in args

    def has_assign_item(self, item=None, dest_buffer_id=None, src_buffer_id=None, index=None,
                        dest_elem_type_id=None, dest_elem_index=None, dest_pointer_offset=None,
                        src_elem_type_id=None, src_elem_index=None, src_pointer_offset=None):
        """python/sixtracklib/trackjob.py: an item given whole, by its eight
        fields, or by (dest_buffer_id, src_buffer_id, index)."""
        j = self.ptr_st_track_job
        if item is not None:
            return bool(st.st_TrackJob_has_assign_address_item(j, ct.byref(self._item_struct(item))))
        if self._detail(dest_elem_type_id, dest_buffer_id, dest_elem_index, dest_pointer_offset,
                        src_elem_type_id, src_buffer_id, src_elem_index, src_pointer_offset):
            return bool(st.st_TrackJob_has_assign_address_item_detailed(
                j, int(dest_elem_type_id), int(dest_buffer_id), int(dest_elem_index),
                int(dest_pointer_offset), int(src_elem_type_id), int(src_buffer_id),
                int(src_elem_index), int(src_pointer_offset)))
        if self._detail(dest_buffer_id, src_buffer_id, index):
            return bool(st.st_TrackJob_has_assign_item_by_index(
                j, int(dest_buffer_id), int(src_buffer_id), int(index)))
        return False

    def index_of_assign_address_item(self, item=None, dest_buffer_id=None, src_buffer_id=None,
                                     dest_elem_type_id=None, dest_elem_index=None,
                                     dest_pointer_offset=None, src_elem_type_id=None,
                                     src_elem_index=None, src_pointer_offset=None):
        """Index of the item among those of its (dest, src) pair; None when
        the job has no such item."""
        j = self.ptr_st_track_job
        if item is not None:
            it = self._item_struct(item)
            if not st.st_TrackJob_has_assign_address_item(j, ct.byref(it)):
                return None
            return int(st.st_TrackJob_index_of_assign_address_item(j, ct.byref(it)))
        if self._detail(dest_elem_type_id, dest_buffer_id, dest_elem_index, dest_pointer_offset,
                        src_elem_type_id, src_buffer_id, src_elem_index, src_pointer_offset):
            args = (int(dest_elem_type_id), int(dest_buffer_id), int(dest_elem_index),
                    int(dest_pointer_offset), int(src_elem_type_id), int(src_buffer_id),
                    int(src_elem_index), int(src_pointer_offset))
            if not st.st_TrackJob_has_assign_address_item_detailed(j, *args):
                return None
            return int(st.st_TrackJob_index_of_assign_address_item_detailed(j, *args))
        return None

    def ptr_assign_address_item(self, item=None, dest_elem_type_id=None, dest_buffer_id=None,
                                dest_elem_index=None, dest_pointer_offset=None,
                                src_elem_type_id=None, src_buffer_id=None, src_elem_index=None,
                                src_pointer_offset=None, index=None):
        """Pointer to the job's copy of an item (st_NullAssignAddressItem
        when there is none)."""
        j = self.ptr_st_track_job
        if item is not None:
            return st.st_TrackJob_ptr_assign_address_item(j, ct.byref(self._item_struct(item)))
        if self._detail(dest_elem_type_id, dest_buffer_id, dest_elem_index, dest_pointer_offset,
                        src_elem_type_id, src_buffer_id, src_elem_index, src_pointer_offset):
            return st.st_TrackJob_ptr_assign_address_item_detailed(
                j, int(dest_elem_type_id), int(dest_buffer_id), int(dest_elem_index),
                int(dest_pointer_offset), int(src_elem_type_id), int(src_buffer_id),
                int(src_elem_index), int(src_pointer_offset))
        if self._detail(dest_buffer_id, src_buffer_id, index):
            return st.st_TrackJob_ptr_assign_address_item_by_index(
                j, int(dest_buffer_id), int(src_buffer_id), int(index))
        return st.st_NullAssignAddressItem

    # -- particle addresses (python/sixtracklib/trackjob.py:1575-1620) ------
    @property
    def can_fetch_particle_addresses(self):
        return bool(st.st_TrackJob_can_fetch_particles_addr(self.ptr_st_track_job))

    @property
    def has_particle_addresses(self):
        return bool(st.st_TrackJob_has_particles_addr(self.ptr_st_track_job))

    def fetch_particle_addresses(self):
        """Device addresses of the fields of every particle set of the job's
        copy (valid until the job is reset or deleted)."""
        self._last_status = st.st_TrackJob_fetch_particle_addresses(self.ptr_st_track_job)
        if self._last_status != 0:
            raise RuntimeError("unsuccessful fetching particle addresses op; status:%d"
                               % self._last_status)
        return self

    def get_particle_addresses(self, particle_set_index=0):
        return st.st_TrackJob_particle_addresses(self.ptr_st_track_job, int(particle_set_index))

    def clear_particle_addresses(self, particle_set_index=0):
        self._last_status = st.st_TrackJob_clear_particle_addresses(self.ptr_st_track_job,
                                                                    int(particle_set_index))
        if self._last_status != 0:
            raise RuntimeError("unsuccessful clearing of particle addresses op; "
                               "particle_set_index=%d, status:%d"
                               % (particle_set_index, self._last_status))
        return self

    def clear_all_particle_addresses(self):
        self._last_status = st.st_TrackJob_clear_all_particle_addresses(self.ptr_st_track_job)
        if self._last_status != 0:
            raise RuntimeError("unsuccessful clearing all particle addresses op; status:%d"
                               % self._last_status)
        return self

    # -- reset, identity, nodes --------------------------------------------
    def reset(self, beam_elements_buffer, particles_buffer, until_turn_elem_by_elem=0,
              output_buffer=None):
        """python/sixtracklib/trackjob.py: the job takes new particle and
        lattice buffers (and an output buffer, made here when one is needed
        and none is given), keeping its configuration."""
        old = self._c_buffers
        self._c_buffers = []
        try:
            eb = get_cbuffer_from_obj(beam_elements_buffer)
            pb = get_cbuffer_from_obj(particles_buffer)
            c_pbuf, c_ebuf = self._map(pb), self._map(eb)
            ob, c_obuf = self._prepare_output(c_pbuf, self._set_indices, c_ebuf,
                                              int(until_turn_elem_by_elem), output_buffer)
            if not st.st_TrackJob_reset_with_output(self.ptr_st_track_job, c_pbuf, c_ebuf, c_obuf,
                                                    int(until_turn_elem_by_elem)):
                raise RuntimeError("unable to reset the TrackJob: " + st.last_error())
        except Exception:
            self._free_buffers()
            self._c_buffers = old
            raise
        for p in old:
            st.st_Buffer_delete(p)
        self._beam_elements_buffer, self._particles_buffer = eb, pb
        self._c_pbuf, self._c_ebuf = c_pbuf, c_ebuf
        self._output_buffer, self._c_obuf = ob, c_obuf
        self._until_turn_ebe = int(until_turn_elem_by_elem)
        return self

    def type(self):
        return int(st.st_TrackJob_get_type_id(self.ptr_st_track_job))

    @property
    def arch_id(self):
        return int(st.st_TrackJob_get_type_id(self.ptr_st_track_job))

    @staticmethod
    def enabled_archs():
        return ["hip"]

    @staticmethod
    def num_all_nodes(arch_str=None):
        """HIP devices for "hip" (or "cuda", the reference's GPU job); 0 for
        other architectures, as the reference gives for disabled ones."""
        if arch_str is not None and str(arch_str).strip().lower() in _GPU_ARCHS + ("cuda",):
            return int(st.st_Hip_num_devices())
        return 0

    @staticmethod
    def num_available_nodes(arch_str=None, env_var_name=None, filter_str=None):
        return TrackJob.num_all_nodes(arch_str)

    @staticmethod
    def print_nodes(arch_str, all=False, filter_str=None, env_var_name=None):
        n = TrackJob.num_all_nodes(arch_str)
        print("%d %s node%s" % (n, str(arch_str).strip().lower(), "" if n == 1 else "s"))
        for k in range(n):
            print("  node 0.%d: HIP device %d" % (k, k))

    # OpenCL programs, kernels, controllers and arguments do not exist for
    # this architecture; the reference answers the same way for every
    # architecture but OpenCL
    @property
    def allows_add_program(self):
        return False

    def add_program(self, path_to_program_file, compile_options):
        raise RuntimeError("Can not add a program to this TrackJob")

    @property
    def allows_enable_kernel(self):
        return False

    def enable_kernel(self, program_id, kernel_name):
        raise RuntimeError("Can not enable a kernel at this TrackJob")

    @property
    def controller(self):
        raise RuntimeError("TrackJob has no controller for this architecture")

    def argument_by_buffer_id(self, buffer_id):
        raise RuntimeError("unable to get argument for buffer on arch %s" % self._arch)

    def stored_buffer_argument(self, buffer_id):
        raise RuntimeError("unable to get argument for stored buffer on arch %s" % self._arch)

    @property
    def opencl_context_addr(self):
        return 0

    @property
    def opencl_queue_addr(self):
        return 0


class TrackJobBaseNew(TrackJob):
    """The new-style job of the reference's Python binding
    (python/sixtracklib/trackjob.py ``TrackJobBaseNew``, driving
    NS(TrackJobNew_*)): built around an existing job pointer, then given its
    buffers by ``_reset_detailed``.  The C job is the same object as the
    legacy job's, so every TrackJob method applies; the reference's
    new-style properties are added below."""

    def __init__(self, ptr_track_job=None, owns_ptr=True):
        self.ptr_st_track_job = ptr_track_job if ptr_track_job else None
        self._owns_ptr = bool(owns_ptr)
        self._c_buffers = []
        self._last_status = 0
        self._last_track_status = 0
        self._arch = "hip"
        self._output_buffer = None
        self._particles_buffer = None
        self._beam_elements_buffer = None
        self._c_pbuf = self._c_ebuf = self._c_obuf = None
        self._until_turn_ebe = 0
        self._set_indices = (st.st_buffer_size_t * 1)(0)
        cfg = (st.st_TrackJobNew_get_config_str(self.ptr_st_track_job)
               if self.ptr_st_track_job else None)
        self._config_str = cfg.decode() if cfg else ""

    def __del__(self, _job_delete=st.st_TrackJob_delete):
        if getattr(self, "ptr_st_track_job", None) and getattr(self, "_owns_ptr", False):
            _job_delete(self.ptr_st_track_job)
        self.ptr_st_track_job = None
        if getattr(self, "_c_buffers", None):
            self._free_buffers()

    def _reset_detailed(self, beam_elements_buffer, particles_buffer, particle_set_index=0,
                        until_turn_elem_by_elem=0, output_buffer=None):
        """Give the job its buffers (one particle set of the particle
        buffer); an output buffer is made when one is needed and none is
        given."""
        if not self.ptr_st_track_job:
            raise ValueError("TrackJob has to be initialized before calling reset")
        old = self._c_buffers
        self._c_buffers = []
        try:
            eb = get_cbuffer_from_obj(beam_elements_buffer)
            pb = get_cbuffer_from_obj(particles_buffer)
            c_pbuf, c_ebuf = self._map(pb), self._map(eb)
            if st.st_Particles_buffer_get_num_of_particle_blocks(c_pbuf) <= particle_set_index:
                raise ValueError("the particle buffer has no particle set %d"
                                 % particle_set_index)
            idx = (st.st_buffer_size_t * 1)(int(particle_set_index))
            ob, c_obuf = self._prepare_output(c_pbuf, idx, c_ebuf, int(until_turn_elem_by_elem),
                                              output_buffer)
            self._last_status = st.st_TrackJobNew_reset_detailed(
                self.ptr_st_track_job, c_pbuf, 1, idx, c_ebuf, c_obuf,
                int(until_turn_elem_by_elem))
            if self._last_status != 0:
                raise RuntimeError("unable to reset the TrackJob: " + st.last_error())
        except Exception:
            self._free_buffers()
            self._c_buffers = old
            raise
        for p in old:
            st.st_Buffer_delete(p)
        self._beam_elements_buffer, self._particles_buffer = eb, pb
        self._c_pbuf, self._c_ebuf = c_pbuf, c_ebuf
        self._output_buffer, self._c_obuf = ob, c_obuf
        self._until_turn_ebe = int(until_turn_elem_by_elem)
        self._set_indices = idx
        return self

    def reset(self, beam_elements_buffer, particles_buffer, until_turn_elem_by_elem=0,
              output_buffer=None):
        return self._reset_detailed(beam_elements_buffer, particles_buffer, 0,
                                    until_turn_elem_by_elem, output_buffer)

    # -- the reference's new-style properties --------------------------------
    @property
    def pointer(self):
        return self.ptr_st_track_job

    @property
    def last_status_success(self):
        return self._last_status == 0

    @property
    def last_track_status_success(self):
        return self._last_track_status == 0

    @property
    def uses_controller(self):
        return bool(st.st_TrackJobNew_uses_controller(self.ptr_st_track_job))

    @property
    def uses_arguments(self):
        return bool(st.st_TrackJobNew_uses_arguments(self.ptr_st_track_job))

    @property
    def is_in_debug_mode(self):
        return bool(st.st_TrackJobNew_is_in_debug_mode(self.ptr_st_track_job))

    @property
    def is_collecting_particles(self):
        return bool(st.st_TrackJobNew_is_collecting_particles(self.ptr_st_track_job))

    @property
    def is_collecting_beam_elements(self):
        return bool(st.st_TrackJobNew_is_collecting_beam_elements(self.ptr_st_track_job))

    @property
    def is_collecting_output(self):
        return bool(st.st_TrackJobNew_is_collecting_output(self.ptr_st_track_job))

    @property
    def collecting_flags(self):
        return int(st.st_TrackJobNew_get_collect_flags(self.ptr_st_track_job))

    @property
    def num_particle_sets(self):
        return int(st.st_TrackJobNew_get_num_particle_sets(self.ptr_st_track_job))

    @property
    def total_num_particles(self):
        return int(st.st_TrackJobNew_get_total_num_of_particles(self.ptr_st_track_job))

    @property
    def has_config_str(self):
        return bool(st.st_TrackJobNew_has_config_str(self.ptr_st_track_job))

    @property
    def owns_output_buffer(self):
        return bool(st.st_TrackJobNew_owns_output_buffer(self.ptr_st_track_job))

    @property
    def has_beam_monitors(self):
        return bool(st.st_TrackJobNew_has_beam_monitors(self.ptr_st_track_job))

    @property
    def beam_monitor_output_buffer_offset(self):
        return int(st.st_TrackJobNew_get_beam_monitor_output_buffer_offset(self.ptr_st_track_job))

    @property
    def elem_by_elem_output_buffer_offset(self):
        return int(st.st_TrackJobNew_get_elem_by_elem_output_buffer_offset(self.ptr_st_track_job))

    @property
    def num_elem_by_elem_turn(self):
        return int(st.st_TrackJobNew_get_num_elem_by_elem_turns(self.ptr_st_track_job))

    @property
    def has_elem_by_elem_config(self):
        return bool(st.st_TrackJobNew_has_elem_by_elem_config(self.ptr_st_track_job))

    def collect(self, collect_flags=None):
        if collect_flags is None:
            st.st_TrackJobNew_collect(self.ptr_st_track_job)
        elif collect_flags > 0:
            st.st_TrackJobNew_collect_detailed(self.ptr_st_track_job, int(collect_flags))
        return self

    def collect_particle_addresses(self):
        self._last_status = st.st_TrackJobNew_collect_particles_addresses(self.ptr_st_track_job)
        return self

    def collect_debug_flag(self):
        self._last_status = st.st_TrackJobNew_collect_debug_flag(self.ptr_st_track_job)
        return self

    def enable_debug_mode(self):
        self._last_status = st.st_TrackJobNew_enable_debug_mode(self.ptr_st_track_job)
        if self._last_status != 0:
            raise RuntimeError("unsuccessful enable debug mode op; status:%d" % self._last_status)
        return self

    def disable_debug_mode(self):
        self._last_status = st.st_TrackJobNew_disable_debug_mode(self.ptr_st_track_job)
        if self._last_status != 0:
            raise RuntimeError("unsuccessful disable debug mode op; status:%d" % self._last_status)
        return self

    # the reference's deprecated camel-case spellings
    def collectParticles(self):
        return self.collect_particles()

    def collectBeamElements(self):
        return self.collect_beam_elements()

    def collectOutput(self):
        return self.collect_output()

    def collectParticlesAddresses(self):
        return self.collect_particle_addresses()

    def collectDebugFlag(self):
        return self.collect_debug_flag()

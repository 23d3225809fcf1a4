"""python/sixtracklib/buffer.py: address-assignment items and the
``Buffer`` handle.

``AssignAddressItem`` is the CObject form of NS(AssignAddressItem)
(buffer/assign_address_item.h:34-45); TrackJob.add_assign_address_item
accepts it, the ctypes structure stcommon.st_AssignAddressItem, or the
eight fields as keywords.  ``Buffer`` owns or borrows an ``NS(Buffer)*``
of the C library: mapped on a CBuffer without a copy, read from a file,
or made empty.
"""
from __future__ import annotations

import ctypes as ct

from . import stcommon as st
from .cobject import CObject, CField
from .trackjob import get_cbuffer_from_obj

__all__ = ["AssignAddressItem", "Buffer", "get_cbuffer_from_obj"]


class AssignAddressItem(CObject):
    _typeid = 513
    dest_buffer_id = CField(0, "uint64", default=0)
    src_buffer_id = CField(1, "uint64", default=0)
    dest_elem_type_id = CField(2, "uint64", default=0)
    src_elem_type_id = CField(3, "uint64", default=0)
    dest_elem_index = CField(4, "uint64", default=1)
    src_elem_index = CField(5, "uint64", default=1)
    dest_pointer_offset = CField(6, "uint64", default=0)
    src_pointer_offset = CField(7, "uint64", default=0)


class Buffer:
    """An ``NS(Buffer)*`` (python/sixtracklib/buffer.py ``Buffer``).

    Exactly one source, in this order: ``cbuffer`` (a CBuffer or an object
    holding one, mapped in place: the C buffer reads and writes the
    CBuffer's memory), ``ptr_ext_buffer`` (an existing pointer, borrowed
    unless ``owns_pointer`` / ``owns_buffer``), ``path_to_file``,
    ``num_objects`` + ``num_slots`` + ``num_dataptrs`` (an empty buffer of
    those capacities), else an empty buffer of ``size`` bytes."""

    def __init__(self, cbuffer=None, ptr_ext_buffer=None, owns_pointer=None, owns_buffer=None,
                 ptr_data=None, path_to_file=None, size=0, num_objects=None, num_slots=None,
                 num_dataptrs=None, num_garbage_ranges=None, buffer_flags=None, slot_size=8):
        self._ptr_st_buffer = None
        self._last_status = 0
        self._owns_buffer = True
        self._owns_pointer = True
        self._cbuffer = None
        if cbuffer is not None:
            cb = get_cbuffer_from_obj(cbuffer)
            self._ptr_st_buffer = st.st_Buffer_new_mapped_on_cbuffer(cb.pointer, cb.words.nbytes)
            self._owns_buffer = False
            if self._ptr_st_buffer:
                self._cbuffer = cb
        elif ptr_ext_buffer:
            self._ptr_st_buffer = ptr_ext_buffer
            self._owns_pointer = owns_pointer is True
            self._owns_buffer = owns_buffer is True
        elif path_to_file:
            self._ptr_st_buffer = st.st_Buffer_new_from_file(str(path_to_file).encode("utf-8"))
        elif ptr_data:
            if not size:
                raise ValueError("ptr_data needs the size of the data in bytes")
            self._ptr_st_buffer = st.st_Buffer_new_on_memory(ptr_data, int(size))
            self._owns_buffer = owns_buffer is True
        elif None not in (num_objects, num_slots, num_dataptrs):
            self._ptr_st_buffer = st.st_Buffer_new_detailed(
                int(num_objects), int(num_slots), int(num_dataptrs),
                int(num_garbage_ranges or 0), int(buffer_flags or 0))
        else:
            self._ptr_st_buffer = st.st_Buffer_new(int(size or 0))
        if not self._ptr_st_buffer:
            raise RuntimeError("unable to create the buffer: " + st.last_error())

    def __del__(self, _delete=st.st_Buffer_delete):
        # a mapped or borrowed buffer's memory is not this object's; the
        # NS(Buffer) handle of a mapping is (st_Buffer_delete leaves the
        # memory it does not own alone)
        if getattr(self, "_ptr_st_buffer", None) and (self._owns_pointer or self._cbuffer is not None):
            _delete(self._ptr_st_buffer)
        self._ptr_st_buffer = None

    @classmethod
    def from_cbuffer(cls, obj, owns_buffer=False, owns_pointer=True, **kwargs):
        cb = get_cbuffer_from_obj(obj)
        return cls(cbuffer=cb, owns_buffer=owns_buffer, owns_pointer=owns_pointer, **kwargs)

    @property
    def pointer(self):
        return self._ptr_st_buffer

    @property
    def maps_to_cbuffer(self):
        return self._cbuffer is not None and bool(self._ptr_st_buffer)

    @property
    def cbuffer(self):
        return self._cbuffer

    @property
    def last_status(self):
        return self._last_status

    @property
    def last_status_success(self):
        return self._last_status == 0

    @property
    def owns_buffer(self):
        return self._owns_buffer

    @property
    def owns_pointer(self):
        return self._owns_pointer

    @property
    def is_owner(self):
        return self._owns_buffer and self._owns_pointer

    @property
    def slot_size(self):
        return int(st.st_Buffer_get_slot_size(self._ptr_st_buffer))

    @property
    def size(self):
        return int(st.st_Buffer_get_size(self._ptr_st_buffer))

    @property
    def capacity(self):
        return int(st.st_Buffer_get_capacity(self._ptr_st_buffer))

    @property
    def num_objects(self):
        return int(st.st_Buffer_get_num_of_objects(self._ptr_st_buffer))

    @property
    def num_slots(self):
        return int(st.st_Buffer_get_num_of_slots(self._ptr_st_buffer))

    @property
    def num_dataptrs(self):
        return int(st.st_Buffer_get_num_of_dataptrs(self._ptr_st_buffer))

    def _header(self):
        """the 8 header words and the data start (managed_buffer_minimal.h)"""
        begin = st.st_Buffer_get_data_begin(self._ptr_st_buffer)
        return (ct.c_uint64 * 8).from_address(begin), begin

    def _section_max(self, k, entity_bytes):
        h, begin = self._header()
        sec = (ct.c_uint64 * 2).from_address(begin + (h[k] - h[0]))
        return int((sec[0] - 16) // entity_bytes)

    @property
    def header_size(self):
        return int(self._header()[0][2])

    @property
    def max_num_slots(self):
        return self._section_max(3, 8)

    @property
    def max_num_objects(self):
        return self._section_max(4, 24)

    @property
    def max_num_dataptrs(self):
        return self._section_max(5, 8)

    @property
    def max_num_garbage_ranges(self):
        return self._section_max(6, 16)

    @property
    def num_garbage_ranges(self):
        h, begin = self._header()
        return int((ct.c_uint64 * 2).from_address(begin + (h[6] - h[0]))[1])

    @property
    def needs_remapping(self):
        return bool(st.st_Buffer_needs_remapping(self._ptr_st_buffer))

    @property
    def data_begin_addr(self):
        return int(st.st_Buffer_get_data_begin_addr(self._ptr_st_buffer))

    @property
    def data_end_addr(self):
        return self.data_begin_addr + self.size

    @property
    def ptr_data_begin(self):
        return st.st_Buffer_get_data_begin(self._ptr_st_buffer)

    @property
    def ptr_data_end(self):
        return self.ptr_data_begin + self.size

    def clear(self, clear_data=False):
        self._last_status = st.st_Buffer_clear(self._ptr_st_buffer, bool(clear_data))
        return self

    def remap(self):
        self._last_status = st.st_Buffer_remap(self._ptr_st_buffer)
        if self._last_status != 0:
            raise RuntimeError("unable to remap the buffer")
        return self

    def write_to_file(self, path_to_file, normalized_base_addr=None):
        if normalized_base_addr is not None:
            raise ValueError("normalised base addresses are not supported")
        ok = st.st_Buffer_write_to_file(self._ptr_st_buffer, str(path_to_file).encode("utf-8"))
        self._last_status = 0 if ok else -1
        return ok

    def reserve(self, max_num_objects=None, max_num_slots=None, max_num_dataptrs=None,
                max_num_garbage_ranges=None):
        """Grow the capacities to at least the given counts (unset: keep)."""
        self._last_status = st.st_Buffer_reserve(
            self._ptr_st_buffer,
            int(max_num_objects if max_num_objects is not None else self.max_num_objects),
            int(max_num_slots if max_num_slots is not None else self.max_num_slots),
            int(max_num_dataptrs if max_num_dataptrs is not None else self.max_num_dataptrs),
            int(max_num_garbage_ranges if max_num_garbage_ranges is not None
                else self.max_num_garbage_ranges))
        if self._last_status != 0:
            raise RuntimeError("unable to reserve buffer capacity")
        return self

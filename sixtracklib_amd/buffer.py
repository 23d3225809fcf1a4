"""Address-assignment items: python/sixtracklib/buffer.py:69-89.

``AssignAddressItem`` is the CObject form of NS(AssignAddressItem)
(buffer/assign_address_item.h:34-45); TrackJob.add_assign_address_item
accepts it, the ctypes structure stcommon.st_AssignAddressItem, or the
eight fields as keywords.
"""
from __future__ import annotations

from .cobject import CObject, CField
from .trackjob import get_cbuffer_from_obj

__all__ = ["AssignAddressItem", "get_cbuffer_from_obj"]


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

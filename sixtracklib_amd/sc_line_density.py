"""Line-density profiles for SCInterpolatedProfile elements:
python/sixtracklib/sc_line_density.py.

A profile (LineDensityProfileData) lives in a buffer the track job keeps
as a stored buffer; an address assignment points an element's
interpol_data_addr at it:

    job = TrackJob(lattice, beam)
    bid = job.add_stored_buffer(cbuffer=profiles)
    LineDensityProfileData_create_buffer_assignment(job, sc_index, bid, 0)
    job.commit_address_assignments()
    job.assign_all_addresses()       # the device lattice now carries the table
"""
from __future__ import annotations

from . import stcommon as st
from .beam_elements import LineDensityProfileData, SCInterpolatedProfile

__all__ = ["LineDensityProfileData", "LineDensityProfileData_create_buffer_assignment"]

# field 6 of NS(SCInterpolatedProfile) (be_beamfields.h:1281-1293)
SC_INTERPOL_DATA_ADDR_OFFSET = 6 * 8


def LineDensityProfileData_create_buffer_assignment(track_job, be_sc_index,
                                                    interpol_buffer_id, lp_data_index):
    """Add the item "interpol_data_addr of element ``be_sc_index`` of the
    beam elements := address of object ``lp_data_index`` of stored buffer
    ``interpol_buffer_id``" (sc_line_density.py:113-152); True when an item
    was added."""
    assert interpol_buffer_id != st.st_ARCH_ILLEGAL_BUFFER_ID.value
    assert track_job.min_stored_buffer_id <= interpol_buffer_id <= track_job.max_stored_buffer_id
    cb = track_job.stored_buffer(interpol_buffer_id)
    assert cb is None or cb.n_objects > lp_data_index
    dest = st.st_ARCH_BEAM_ELEMENTS_BUFFER_ID.value
    before = track_job.num_assign_items(dest, interpol_buffer_id)
    track_job.add_assign_address_item(
        dest_elem_type_id=SCInterpolatedProfile._typeid, dest_buffer_id=dest,
        dest_elem_index=be_sc_index, dest_pointer_offset=SC_INTERPOL_DATA_ADDR_OFFSET,
        src_elem_type_id=LineDensityProfileData._typeid, src_buffer_id=interpol_buffer_id,
        src_elem_index=lp_data_index, src_pointer_offset=LineDensityProfileData.ptr_offset())
    return before < track_job.num_assign_items(dest, interpol_buffer_id)

"""ctypes bindings to libsixtrack_amd.so.

Counterpart of python/sixtracklib/stcommon.py (``ct.CDLL(stconf.SHARED_LIB)``
and the ``st_*`` prototypes, e.g. :1334-1375 for the track job).  The
library is loaded from the package directory; if it is missing this module
raises ImportError -- there is no pure-Python fallback for tracking.
"""
from __future__ import annotations

import ctypes as ct
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SHARED_LIB = os.path.join(HERE, "libsixtrack_amd.so")

if not os.path.exists(SHARED_LIB):
    raise ImportError("libsixtrack_amd.so not built: run `python -m sixtracklib_amd.build` "
                      "(or __graft_entry__.build())")


def hip_runtime_path():
    """Path of PyTorch's HIP runtime (libamdhip64.so next to libtorch_hip.so)
    when torch is installed, else None."""
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    return path if os.path.exists(path) else None


def _share_torch_hip_runtime():
    """PyTorch-ROCm wheels carry their own HIP runtime, which torch loads
    by file name (libamdhip64.so next to libtorch_hip.so).  Loaded first,
    this library would bind the system runtime (soname libamdhip64.so.7),
    a later ``import torch`` would bring a second runtime into the process
    and torch would then find no GPU.  Loading torch's copy first (without
    importing torch) gives the process one runtime, shared by both.
    SIXTRACKLIB_AMD_SYSTEM_HIP=1 skips this."""
    if os.environ.get("SIXTRACKLIB_AMD_SYSTEM_HIP") == "1":
        return
    path = hip_runtime_path()
    if path is not None:
        ct.CDLL(path, mode=ct.RTLD_GLOBAL)


_share_torch_hip_runtime()
sixtracklib = ct.CDLL(SHARED_LIB)

st_buffer_size_t = ct.c_uint64
st_buffer_addr_t = ct.c_uint64
st_track_status_t = ct.c_int32
st_particle_index_t = ct.c_int64
st_Buffer_p = ct.c_void_p
st_NullBuffer = None
st_TrackJob_p = ct.c_void_p
st_NullTrackJob = None
st_uchar_p = ct.c_void_p
st_buffer_size_p = ct.POINTER(st_buffer_size_t)

st_TRACK_SUCCESS = 0
st_TRACK_STATUS_GENERAL_FAILURE = -1
st_TRACK_JOB_IO_PARTICLES = 0x01
st_TRACK_JOB_IO_BEAM_ELEMENTS = 0x02
st_TRACK_JOB_IO_OUTPUT = 0x04


def _fn(name, restype, argtypes):
    f = getattr(sixtracklib, name)
    f.restype = restype
    f.argtypes = argtypes
    globals()[name] = f
    return f


# -- buffers ------------------------------------------------------------
_fn("st_Buffer_new", st_Buffer_p, [st_buffer_size_t])
_fn("st_Buffer_new_detailed", st_Buffer_p, [st_buffer_size_t] * 4 + [ct.c_uint64])
_fn("st_Buffer_new_from_file", st_Buffer_p, [ct.c_char_p])
_fn("st_Buffer_new_on_memory", st_Buffer_p, [st_uchar_p, st_buffer_size_t])
_fn("st_Buffer_new_mapped_on_cbuffer", st_Buffer_p, [st_uchar_p, st_buffer_size_t])
_fn("st_Buffer_new_from_copy", st_Buffer_p, [st_Buffer_p])
_fn("st_Buffer_delete", None, [st_Buffer_p])
_fn("st_Buffer_write_to_file", ct.c_bool, [st_Buffer_p, ct.c_char_p])
_fn("st_Buffer_needs_remapping", ct.c_bool, [st_Buffer_p])
_fn("st_Buffer_remap", ct.c_int, [st_Buffer_p])
_fn("st_Buffer_get_num_of_objects", st_buffer_size_t, [st_Buffer_p])
_fn("st_Buffer_get_size", st_buffer_size_t, [st_Buffer_p])
_fn("st_Buffer_get_capacity", st_buffer_size_t, [st_Buffer_p])
_fn("st_Buffer_get_data_begin_addr", st_buffer_addr_t, [st_Buffer_p])
_fn("st_Buffer_get_object", ct.c_void_p, [st_Buffer_p, st_buffer_size_t])
_fn("st_ManagedBuffer_remap_to", ct.c_int, [st_uchar_p, st_buffer_addr_t])

# -- particles / elements -------------------------------------------------
_fn("st_Particles_new", ct.c_void_p, [st_Buffer_p, ct.c_int64])
_fn("st_Particles_buffer_get_particles", ct.c_void_p, [st_Buffer_p, st_buffer_size_t])
_fn("st_Particles_buffer_get_num_of_particle_blocks", st_buffer_size_t, [st_Buffer_p])
_fn("st_Particles_buffer_get_total_num_of_particles", ct.c_int64, [st_Buffer_p])
_fn("st_Drift_add", ct.c_void_p, [st_Buffer_p, ct.c_double])
_fn("st_DriftExact_add", ct.c_void_p, [st_Buffer_p, ct.c_double])
_fn("st_Multipole_add", ct.c_void_p, [st_Buffer_p, ct.c_int64, ct.c_double,
                                      ct.c_double, ct.c_double, st_buffer_addr_t])
_fn("st_Cavity_add", ct.c_void_p, [st_Buffer_p, ct.c_double, ct.c_double, ct.c_double])
_fn("st_LimitRect_add", ct.c_void_p, [st_Buffer_p] + [ct.c_double] * 4)
_fn("st_Math_q_gauss_cq", ct.c_double, [ct.c_double])
_fn("st_Lattice_pack", ct.c_int64, [st_uchar_p, ct.c_void_p, ct.c_int64, ct.c_void_p,
                                    ct.POINTER(ct.c_int64), ct.c_void_p,
                                    ct.POINTER(ct.c_int64)])

# -- track job ----------------------------------------------------------
_fn("st_TrackJob_create", st_TrackJob_p, [ct.c_char_p, ct.c_char_p])
_fn("st_TrackJob_new", st_TrackJob_p, [ct.c_char_p, st_Buffer_p, st_Buffer_p, ct.c_char_p])
_fn("st_TrackJob_new_with_output", st_TrackJob_p,
    [ct.c_char_p, st_Buffer_p, st_Buffer_p, st_Buffer_p, st_buffer_size_t, ct.c_char_p])
_fn("st_TrackJob_new_detailed", st_TrackJob_p,
    [ct.c_char_p, st_Buffer_p, st_buffer_size_t, st_buffer_size_p, st_Buffer_p,
     st_Buffer_p, st_buffer_size_t, ct.c_char_p])
_fn("st_TrackJob_delete", None, [st_TrackJob_p])
_fn("st_TrackJob_track_until", st_track_status_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_track_until_noalloc", st_track_status_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_track_elem_by_elem", st_track_status_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_track_line", st_track_status_t,
    [st_TrackJob_p, st_buffer_size_t, st_buffer_size_t, ct.c_bool])
for _n in ("collect", "collect_particles", "collect_beam_elements", "collect_output",
           "push_particles", "push_beam_elements", "push_output", "clear",
           "enable_collect_particles", "disable_collect_particles",
           "enable_collect_output", "disable_collect_output"):
    _fn("st_TrackJob_" + _n, None, [st_TrackJob_p])
_fn("st_TrackJob_collect_detailed", None, [st_TrackJob_p, ct.c_uint16])
_fn("st_TrackJob_push", None, [st_TrackJob_p, ct.c_uint16])
for _n in ("has_output_buffer", "owns_output_buffer", "has_elem_by_elem_output",
           "has_beam_monitor_output", "has_device_id_str", "has_config_str",
           "requires_collecting", "is_collecting_particles", "is_collecting_output"):
    _fn("st_TrackJob_" + _n, ct.c_bool, [st_TrackJob_p])
for _n in ("get_beam_monitor_output_buffer_offset", "get_elem_by_elem_output_buffer_offset",
           "get_num_elem_by_elem_turns", "get_num_beam_monitors", "get_num_particle_sets",
           "get_total_num_particles_in_sets"):
    _fn("st_TrackJob_" + _n, st_buffer_size_t, [st_TrackJob_p])
for _n in ("get_min_particle_id", "get_max_particle_id", "get_min_element_id",
           "get_max_element_id", "get_min_initial_turn_id", "get_max_initial_turn_id"):
    _fn("st_TrackJob_" + _n, st_particle_index_t, [st_TrackJob_p])
_fn("st_TrackJob_get_type_str", ct.c_char_p, [st_TrackJob_p])
_fn("st_TrackJob_get_config_str", ct.c_char_p, [st_TrackJob_p])
_fn("st_TrackJob_get_device_id_str", ct.c_char_p, [st_TrackJob_p])
_fn("st_TrackJob_get_output_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_last_status", st_track_status_t, [st_TrackJob_p])
_fn("st_TrackJob_required_output_sets", ct.c_int64,
    [st_Buffer_p, st_buffer_size_t, st_buffer_size_p, st_Buffer_p, st_buffer_size_t,
     ct.POINTER(ct.c_int64), ct.c_int64, ct.POINTER(ct.c_int64), ct.c_char_p])
_fn("st_TrackJob_last_kernel_time_ms", ct.c_double, [st_TrackJob_p])
_fn("st_TrackJob_uses_fma", ct.c_int, [st_TrackJob_p])
_fn("st_TrackJob_uses_full_kernel", ct.c_int, [st_TrackJob_p])

# -- batched multi-seed tracking (include/sixtracklib_amd/track_batch.h) ---
st_TrackBatch_p = ct.c_void_p
_fn("st_TrackBatch_new", st_TrackBatch_p,
    [ct.c_int64, ct.POINTER(st_Buffer_p), ct.POINTER(st_Buffer_p), ct.c_char_p])
_fn("st_TrackBatch_delete", None, [st_TrackBatch_p])
_fn("st_TrackBatch_track_until", st_track_status_t, [st_TrackBatch_p, ct.c_int64])
_fn("st_TrackBatch_collect", ct.c_int, [st_TrackBatch_p])
_fn("st_TrackBatch_push", ct.c_int, [st_TrackBatch_p])
_fn("st_TrackBatch_set_config", ct.c_int, [st_TrackBatch_p, ct.c_char_p])
_fn("st_TrackBatch_get_config", ct.c_int64, [st_TrackBatch_p, ct.c_char_p, ct.c_int64])
_fn("st_TrackBatch_autotune", ct.c_int, [st_TrackBatch_p, ct.c_int64])
_fn("st_TrackBatch_autotune_window", ct.c_int, [st_TrackBatch_p, ct.c_int64, ct.c_int64])
_fn("st_TrackBatch_num_seeds", ct.c_int64, [st_TrackBatch_p])
_fn("st_TrackBatch_total_particles", ct.c_int64, [st_TrackBatch_p])
_fn("st_TrackBatch_last_kernel_time_ms", ct.c_double, [st_TrackBatch_p])

# -- stored buffers and address assignments (track_job.h:500-772) ----------
st_arch_size_t = ct.c_uint64
st_arch_status_t = ct.c_int32
st_object_type_id_t = ct.c_uint64
st_ARCH_STATUS_SUCCESS = ct.c_int32(0)
st_ARCH_STATUS_GENERAL_FAILURE = ct.c_int32(-1)
st_ARCH_ILLEGAL_BUFFER_ID = ct.c_uint64(0)
st_ARCH_PARTICLES_BUFFER_ID = ct.c_uint64(1)
st_ARCH_BEAM_ELEMENTS_BUFFER_ID = ct.c_uint64(2)
st_ARCH_OUTPUT_BUFFER_ID = ct.c_uint64(3)
st_ARCH_MIN_USER_DEFINED_BUFFER_ID = ct.c_uint64(1024)


class st_AssignAddressItem(ct.Structure):
    """buffer/assign_address_item.h:34-45"""
    _fields_ = [(n, ct.c_uint64) for n in (
        "dest_buffer_id", "src_buffer_id", "dest_elem_type_id", "src_elem_type_id",
        "dest_elem_index", "src_elem_index", "dest_pointer_offset", "src_pointer_offset")]


st_AssignAddressItem_p = ct.POINTER(st_AssignAddressItem)
st_NullAssignAddressItem = None

_fn("st_TrackJob_add_stored_buffer", st_arch_size_t,
    [st_TrackJob_p, st_Buffer_p, ct.c_bool, ct.c_bool])
_fn("st_TrackJob_create_stored_buffer", st_arch_size_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_remove_stored_buffer", st_arch_status_t, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_has_stored_buffers", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_num_stored_buffers", st_arch_size_t, [st_TrackJob_p])
_fn("st_TrackJob_min_stored_buffer_id", st_arch_size_t, [st_TrackJob_p])
_fn("st_TrackJob_max_stored_buffer_id", st_arch_size_t, [st_TrackJob_p])
_fn("st_TrackJob_owns_stored_buffer", ct.c_bool, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_stored_buffer", st_Buffer_p, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_push_stored_buffer", st_arch_status_t, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_collect_stored_buffer", st_arch_status_t, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_add_assign_address_item", st_AssignAddressItem_p,
    [st_TrackJob_p, st_AssignAddressItem_p])
_fn("st_TrackJob_add_assign_address_item_detailed", st_AssignAddressItem_p,
    [st_TrackJob_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJob_has_assign_items", ct.c_bool, [st_TrackJob_p, st_arch_size_t, st_arch_size_t])
_fn("st_TrackJob_num_assign_items", st_arch_size_t,
    [st_TrackJob_p, st_arch_size_t, st_arch_size_t])
_fn("st_TrackJob_total_num_assign_items", st_arch_size_t, [st_TrackJob_p])
_fn("st_TrackJob_commit_address_assignments", st_arch_status_t, [st_TrackJob_p])
_fn("st_TrackJob_assign_all_addresses", st_arch_status_t, [st_TrackJob_p])
_fn("st_TrackJob_assign_addresses", st_arch_status_t,
    [st_TrackJob_p, st_buffer_size_t, st_buffer_size_t])
_fn("st_Lattice_pack_with_tables", ct.c_int64,
    [st_uchar_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
     ct.c_void_p, ct.c_void_p])

# -- the rest of the legacy job API (track_job.h:123-470) -----------------
class st_ParticlesAddr(ct.Structure):
    """particles/particles_addr.h:27-55"""
    _fields_ = [("num_particles", ct.c_int64)] + [(n + "_addr", ct.c_uint64) for n in (
        "q0", "mass0", "beta0", "gamma0", "p0c", "s", "x", "y", "px", "py", "zeta", "psigma",
        "delta", "rpp", "rvv", "chi", "charge_ratio", "particle_id", "at_element_id", "at_turn",
        "state")]


class st_ElemByElemConfig(ct.Structure):
    """output/elem_by_elem_config.h:39-53"""
    _fields_ = [(n, ct.c_int64) for n in (
        "order", "num_particles_to_store", "num_elements_to_store", "num_turns_to_store",
        "min_particle_id", "min_element_id", "min_turn", "max_particle_id", "max_element_id",
        "max_turn", "is_rolling")] + [("out_store_addr", ct.c_uint64)]


class st_AssignAddressItemDestSrcPair(ct.Structure):
    _fields_ = [("dest_buffer_id", ct.c_uint64), ("src_buffer_id", ct.c_uint64)]


st_ParticlesAddr_p = ct.POINTER(st_ParticlesAddr)
_fn("st_TrackJob_extract_device_id_str", ct.c_int, [ct.c_char_p, ct.c_char_p, st_buffer_size_t])
_fn("st_TrackJob_sanitize_arch_str", ct.c_int, [ct.c_char_p, ct.c_char_p, st_buffer_size_t])
_fn("st_TrackJob_sanitize_arch_str_inplace", ct.c_int, [ct.c_char_p, st_buffer_size_t])
_fn("st_TrackJob_can_fetch_particles_addr", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_has_particles_addr", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_fetch_particle_addresses", st_arch_status_t, [st_TrackJob_p])
_fn("st_TrackJob_clear_all_particle_addresses", st_arch_status_t, [st_TrackJob_p])
_fn("st_TrackJob_particle_addresses", st_ParticlesAddr_p, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_get_elem_by_elem_config", ct.POINTER(st_ElemByElemConfig), [st_TrackJob_p])
_fn("st_TrackJob_get_beam_monitor_indices_begin", st_buffer_size_p, [st_TrackJob_p])
_fn("st_TrackJob_get_num_particles_in_set", st_buffer_size_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_block_size", ct.c_int, [st_TrackJob_p])
_fn("st_TrackJob_remove_assign_address_item", st_arch_status_t,
    [st_TrackJob_p, st_AssignAddressItem_p])
_fn("st_TrackJob_index_of_assign_address_item", st_buffer_size_t,
    [st_TrackJob_p, st_AssignAddressItem_p])
_fn("st_TrackJob_ptr_assign_address_item_by_index", st_AssignAddressItem_p,
    [st_TrackJob_p, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJob_num_distinct_available_assign_address_items_dest_src_pairs", st_buffer_size_t,
    [st_TrackJob_p])
_fn("st_TrackJob_available_assign_address_items_dest_src_pairs", st_buffer_size_t,
    [st_TrackJob_p, st_buffer_size_t, ct.POINTER(st_AssignAddressItemDestSrcPair)])
_fn("st_TrackJob_assign_items_dest_src_begin", st_AssignAddressItem_p,
    [st_TrackJob_p, st_arch_size_t, st_arch_size_t])
_fn("st_TrackJob_assign_items_dest_src_end", st_AssignAddressItem_p,
    [st_TrackJob_p, st_arch_size_t, st_arch_size_t])
_fn("st_TrackJob_assign_items_begin", ct.c_void_p, [st_TrackJob_p, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJob_assign_items_end", ct.c_void_p, [st_TrackJob_p, st_buffer_size_t, st_buffer_size_t])

# -- new-style job API and the CUDA job's constructors (track_job_new.h) ---
st_TrackJobNew_p = ct.c_void_p
st_ARCHITECTURE_HIP = 4
_fn("st_TrackJobNew_create", st_TrackJobNew_p, [ct.c_char_p, ct.c_char_p])
_fn("st_TrackJobNew_new", st_TrackJobNew_p, [ct.c_char_p, st_Buffer_p, st_Buffer_p, ct.c_char_p])
_fn("st_TrackJobNew_new_with_output", st_TrackJobNew_p,
    [ct.c_char_p, st_Buffer_p, st_Buffer_p, st_Buffer_p, st_buffer_size_t, ct.c_char_p])
_fn("st_TrackJobNew_new_detailed", st_TrackJobNew_p,
    [ct.c_char_p, st_Buffer_p, st_buffer_size_t, st_buffer_size_p, st_Buffer_p, st_Buffer_p,
     st_buffer_size_t, ct.c_char_p])
_fn("st_TrackJobNew_delete", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_track_until", st_track_status_t, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_track_elem_by_elem", st_track_status_t, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_track_line", st_track_status_t,
    [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t, ct.c_bool])
_fn("st_TrackJobNew_collect", ct.c_uint16, [st_TrackJobNew_p])
_fn("st_TrackJobNew_collect_particles", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_collect_output", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_push_particles", ct.c_uint16, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_arch_id", ct.c_uint64, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_arch_string", ct.c_char_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_uses_controller", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_uses_arguments", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_is_in_debug_mode", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_enable_debug_mode", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_output_buffer", st_Buffer_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_elem_by_elem_output_buffer_offset", st_buffer_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_total_num_of_particles", st_buffer_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_fetch_particle_addresses", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_particle_addresses", st_ParticlesAddr_p, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_add_assign_address_item_detailed", st_AssignAddressItem_p,
    [st_TrackJobNew_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_total_num_assign_items", st_buffer_size_t, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_num_available_nodes", ct.c_uint64, [])
_fn("st_CudaTrackJob_get_available_node_indices_list", ct.c_uint64,
    [ct.c_uint64, ct.POINTER(ct.c_uint32)])
_fn("st_CudaTrackJob_create", st_TrackJobNew_p, [])
_fn("st_CudaTrackJob_new", st_TrackJobNew_p, [ct.c_char_p, st_Buffer_p, st_Buffer_p])
_fn("st_CudaTrackJob_new_with_output", st_TrackJobNew_p,
    [ct.c_char_p, st_Buffer_p, st_Buffer_p, st_Buffer_p, st_buffer_size_t])
_fn("st_CudaTrackJob_new_detailed", st_TrackJobNew_p,
    [ct.c_char_p, st_Buffer_p, st_buffer_size_t, st_buffer_size_p, st_Buffer_p, st_Buffer_p,
     st_buffer_size_t, ct.c_char_p])
_fn("st_CudaTrackJob_has_controller", ct.c_bool, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_particles_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_default_threads_per_block", ct.c_uint64, [st_TrackJobNew_p])

# -- the rest of the C API of the headers above -----------------------------
# (buffer.h, particles.h, beam_elements.h, track_job.h, track_job_new.h; the
# adjoint's entry points are bound in adjoint.py with its problem struct)


class st_NodeId(ct.Structure):
    """NS(NodeId) of track_job_new.h: a HIP device as a node"""
    _fields_ = [("platform_id", ct.c_int64), ("device_id", ct.c_int64),
                ("node_index", ct.c_uint32)]


_fn("st_Buffer_clear", ct.c_int, [st_Buffer_p, ct.c_bool])
_fn("st_Buffer_reserve", ct.c_int,
    [st_Buffer_p, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_Buffer_get_slot_size", st_buffer_size_t, [st_Buffer_p])
_fn("st_Buffer_get_num_of_slots", st_buffer_size_t, [st_Buffer_p])
_fn("st_Buffer_get_num_of_dataptrs", st_buffer_size_t, [st_Buffer_p])
_fn("st_Buffer_get_data_begin", st_uchar_p, [st_Buffer_p])
_fn("st_Buffer_get_const_data_begin", st_uchar_p, [st_Buffer_p])
_fn("st_Buffer_owns_datastore", ct.c_bool, [st_Buffer_p])
_fn("st_Buffer_get_objects_begin", ct.c_void_p, [st_Buffer_p])
_fn("st_Buffer_get_objects_end", ct.c_void_p, [st_Buffer_p])
_fn("st_Buffer_get_const_object", ct.c_void_p, [st_Buffer_p, st_buffer_size_t])
_fn("st_Buffer_add_object", ct.c_void_p,
    [st_Buffer_p, ct.c_void_p, st_buffer_size_t, st_object_type_id_t, st_buffer_size_t, st_buffer_size_p,
     st_buffer_size_p, st_buffer_size_p])
_fn("st_Object_get_type_id", st_object_type_id_t, [ct.c_void_p])
_fn("st_Object_get_begin_addr", st_buffer_addr_t, [ct.c_void_p])
_fn("st_Object_get_size", st_buffer_size_t, [ct.c_void_p])
_fn("st_Particles_add_copy", ct.c_void_p, [st_Buffer_p, ct.c_void_p])
_fn("st_Particles_buffer_get_const_particles", ct.c_void_p, [st_Buffer_p, st_buffer_size_t])
_fn("st_Particles_get_num_of_particles", ct.c_int64, [ct.c_void_p])
_fn("st_Particles_copy", ct.c_int, [ct.c_void_p, ct.c_void_p])
_fn("st_Particles_find_min_max_attributes", ct.c_int,
    [ct.c_void_p, ct.POINTER(st_particle_index_t), ct.POINTER(st_particle_index_t),
     ct.POINTER(st_particle_index_t), ct.POINTER(st_particle_index_t), ct.POINTER(st_particle_index_t),
     ct.POINTER(st_particle_index_t)])
_fn("st_Multipole_new", ct.c_void_p, [st_Buffer_p, ct.c_int64])
_fn("st_RFMultipole_add", ct.c_void_p,
    [st_Buffer_p, ct.c_int64, ct.c_double, ct.c_double, ct.c_double, st_buffer_addr_t, st_buffer_addr_t])
_fn("st_XYShift_add", ct.c_void_p, [st_Buffer_p, ct.c_double, ct.c_double])
_fn("st_SRotation_add", ct.c_void_p, [st_Buffer_p, ct.c_double, ct.c_double])
_fn("st_LimitEllipse_add", ct.c_void_p, [st_Buffer_p, ct.c_double, ct.c_double])
_fn("st_LimitRectEllipse_add", ct.c_void_p, [st_Buffer_p, ct.c_double, ct.c_double, ct.c_double, ct.c_double])
_fn("st_DipoleEdge_add", ct.c_void_p, [st_Buffer_p, ct.c_double, ct.c_double])
_fn("st_BeamMonitor_add", ct.c_void_p,
    [st_Buffer_p, ct.c_int64, ct.c_int64, ct.c_int64, st_buffer_addr_t, ct.c_int64, ct.c_int64, ct.c_bool,
     ct.c_bool])
_fn("st_BeamBeam4D_add", ct.c_void_p, [st_Buffer_p, ct.POINTER(ct.c_double), ct.c_uint64])
_fn("st_BeamBeam6D_add", ct.c_void_p, [st_Buffer_p, ct.POINTER(ct.c_double), ct.c_uint64])
_fn("st_SCCoasting_add", ct.c_void_p,
    [st_Buffer_p, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_double,
     ct.c_double, ct.c_bool])
_fn("st_SCQGaussProfile_add", ct.c_void_p,
    [st_Buffer_p, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_double,
     ct.c_double, ct.c_double, ct.c_bool])
_fn("st_TrackJob_enable_collect_beam_elements", None, [st_TrackJob_p])
_fn("st_TrackJob_disable_collect_beam_elements", None, [st_TrackJob_p])
_fn("st_TrackJob_is_collecting_beam_elements", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_get_collect_flags", ct.c_uint16, [st_TrackJob_p])
_fn("st_TrackJob_set_collect_flags", None, [st_TrackJob_p, ct.c_uint16])
_fn("st_TrackJob_collect_particles_addresses", None, [st_TrackJob_p])
_fn("st_TrackJob_push_particles_addresses", None, [st_TrackJob_p])
_fn("st_TrackJob_fetch_particles_addr", st_arch_status_t, [st_TrackJob_p])
_fn("st_TrackJob_clear_particle_addresses", st_arch_status_t, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_clear_particles_addr", st_arch_status_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_clear_all_particles_addr", st_arch_status_t, [st_TrackJob_p])
_fn("st_TrackJob_particles_addr", st_ParticlesAddr_p, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_get_particles_addr_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_get_const_particles_addr_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_particles_addr_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_reset", ct.c_bool, [st_TrackJob_p, st_Buffer_p, st_Buffer_p, st_Buffer_p])
_fn("st_TrackJob_reset_with_output", ct.c_bool,
    [st_TrackJob_p, st_Buffer_p, st_Buffer_p, st_Buffer_p, st_buffer_size_t])
_fn("st_TrackJob_reset_detailed", ct.c_bool,
    [st_TrackJob_p, st_Buffer_p, st_buffer_size_t, st_buffer_size_p, st_Buffer_p, st_Buffer_p,
     st_buffer_size_t])
_fn("st_TrackJob_reset_particle_set", ct.c_bool,
    [st_TrackJob_p, st_Buffer_p, st_buffer_size_t, st_Buffer_p, st_Buffer_p])
_fn("st_TrackJob_select_particle_set", ct.c_bool, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_assign_output_buffer", ct.c_bool, [st_TrackJob_p, st_Buffer_p])
_fn("st_TrackJob_get_type_id", ct.c_int64, [st_TrackJob_p])
_fn("st_TrackJob_get_particle_set_indices_begin", st_buffer_size_p, [st_TrackJob_p])
_fn("st_TrackJob_get_particle_set_indices_end", st_buffer_size_p, [st_TrackJob_p])
_fn("st_TrackJob_get_particle_set_index", st_buffer_size_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_get_num_particles_in_sets_begin", st_buffer_size_p, [st_TrackJob_p])
_fn("st_TrackJob_get_num_particles_in_sets_end", st_buffer_size_p, [st_TrackJob_p])
_fn("st_TrackJob_get_particles_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_get_const_particles_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_get_beam_elements_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_get_const_beam_elements_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_get_const_output_buffer", st_Buffer_p, [st_TrackJob_p])
_fn("st_TrackJob_has_beam_monitors", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_get_beam_monitor_indices_end", st_buffer_size_p, [st_TrackJob_p])
_fn("st_TrackJob_get_beam_monitor_index", st_buffer_size_t, [st_TrackJob_p, st_buffer_size_t])
_fn("st_TrackJob_has_elem_by_elem_config", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_is_elem_by_elem_config_rolling", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_get_default_elem_by_elem_config_rolling_flag", ct.c_bool, [st_TrackJob_p])
_fn("st_TrackJob_set_default_elem_by_elem_config_rolling_flag", None, [st_TrackJob_p, ct.c_bool])
_fn("st_TrackJob_get_elem_by_elem_config_order", ct.c_int64, [st_TrackJob_p])
_fn("st_TrackJob_get_default_elem_by_elem_config_order", ct.c_int64, [st_TrackJob_p])
_fn("st_TrackJob_set_default_elem_by_elem_config_order", None, [st_TrackJob_p, ct.c_int64])
_fn("st_TrackJob_const_stored_buffer", st_Buffer_p, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_stored_buffers_capacity", st_arch_size_t, [st_TrackJob_p])
_fn("st_TrackJob_reserve_stored_buffers_capacity", st_arch_status_t, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_buffer_by_buffer_id", st_Buffer_p, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_const_buffer_by_buffer_id", st_Buffer_p, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_is_buffer_by_buffer_id", ct.c_bool, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_is_raw_memory_by_buffer_id", ct.c_bool, [st_TrackJob_p, st_arch_size_t])
_fn("st_TrackJob_add_assign_address_item_detail", st_AssignAddressItem_p,
    [st_TrackJob_p, st_object_type_id_t, st_arch_size_t, st_arch_size_t, st_arch_size_t, st_object_type_id_t,
     st_arch_size_t, st_arch_size_t, st_arch_size_t])
_fn("st_TrackJob_remove_assign_address_item_by_key_and_index", st_arch_status_t,
    [st_TrackJob_p, ct.POINTER(st_AssignAddressItemDestSrcPair), st_buffer_size_t])
_fn("st_TrackJob_has_assign_address_item", ct.c_bool, [st_TrackJob_p, st_AssignAddressItem_p])
_fn("st_TrackJob_has_assign_item_by_index", ct.c_bool,
    [st_TrackJob_p, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJob_has_assign_address_item_detailed", ct.c_bool,
    [st_TrackJob_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJob_index_of_assign_address_item_detailed", st_buffer_size_t,
    [st_TrackJob_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJob_ptr_assign_address_item", st_AssignAddressItem_p, [st_TrackJob_p, st_AssignAddressItem_p])
_fn("st_TrackJob_ptr_assign_address_item_detailed", st_AssignAddressItem_p,
    [st_TrackJob_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJob_num_distinct_assign_items_dest_src_pairs", st_arch_size_t, [st_TrackJob_p])
_fn("st_TrackJob_assign_items_dest_src_pairs", st_arch_size_t,
    [st_TrackJob_p, st_arch_size_t, ct.POINTER(st_AssignAddressItemDestSrcPair)])
_fn("st_TrackJob_assign_item_dest_src_begin", ct.POINTER(st_AssignAddressItemDestSrcPair), [st_TrackJob_p])
_fn("st_TrackJob_assign_item_dest_src_end", ct.POINTER(st_AssignAddressItemDestSrcPair), [st_TrackJob_p])
_fn("st_TrackJobNew_collect_detailed", ct.c_uint16, [st_TrackJobNew_p, ct.c_uint16])
_fn("st_TrackJobNew_collect_beam_elements", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_collect_debug_flag", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_collect_particles_addresses", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_enable_collect_particles", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_disable_collect_particles", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_is_collecting_particles", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_enable_collect_beam_elements", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_disable_collect_beam_elements", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_is_collecting_beam_elements", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_enable_collect_output", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_disable_collect_output", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_is_collecting_output", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_collect_flags", ct.c_uint16, [st_TrackJobNew_p])
_fn("st_TrackJobNew_set_collect_flags", None, [st_TrackJobNew_p, ct.c_uint16])
_fn("st_TrackJobNew_requires_collecting", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_push", ct.c_uint16, [st_TrackJobNew_p, ct.c_uint16])
_fn("st_TrackJobNew_push_beam_elements", ct.c_uint16, [st_TrackJobNew_p])
_fn("st_TrackJobNew_push_output", ct.c_uint16, [st_TrackJobNew_p])
_fn("st_TrackJobNew_can_fetch_particle_addresses", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_has_particle_addresses", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_clear_particle_addresses", st_arch_status_t, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_clear_all_particle_addresses", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_ptr_particle_addresses_buffer", st_Buffer_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_disable_debug_mode", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_clear", None, [st_TrackJobNew_p])
_fn("st_TrackJobNew_reset", st_arch_status_t, [st_TrackJobNew_p, st_Buffer_p, st_Buffer_p, st_Buffer_p])
_fn("st_TrackJobNew_reset_particle_set", st_arch_status_t,
    [st_TrackJobNew_p, st_Buffer_p, st_buffer_size_t, st_Buffer_p, st_Buffer_p])
_fn("st_TrackJobNew_reset_with_output", st_arch_status_t,
    [st_TrackJobNew_p, st_Buffer_p, st_Buffer_p, st_Buffer_p, st_buffer_size_t])
_fn("st_TrackJobNew_reset_detailed", st_arch_status_t,
    [st_TrackJobNew_p, st_Buffer_p, st_buffer_size_t, st_buffer_size_p, st_Buffer_p, st_Buffer_p,
     st_buffer_size_t])
_fn("st_TrackJobNew_select_particle_set", st_arch_status_t, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_assign_output_buffer", st_arch_status_t, [st_TrackJobNew_p, st_Buffer_p])
_fn("st_TrackJobNew_has_arch_string", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_has_config_str", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_config_str", ct.c_char_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_num_particle_sets", st_buffer_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_particle_set_indices_begin", st_buffer_size_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_particle_set_indices_end", st_buffer_size_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_particle_set_index", st_buffer_size_t, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_get_min_particle_id", st_particle_index_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_max_particle_id", st_particle_index_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_min_element_id", st_particle_index_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_max_element_id", st_particle_index_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_min_initial_turn_id", st_particle_index_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_max_initial_turn_id", st_particle_index_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_particles_buffer", st_Buffer_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_const_particles_buffer", st_Buffer_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_beam_elements_buffer", st_Buffer_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_const_beam_elements_buffer", st_Buffer_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_has_output_buffer", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_owns_output_buffer", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_has_elem_by_elem_output", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_has_beam_monitor_output", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_beam_monitor_output_buffer_offset", st_buffer_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_num_elem_by_elem_turns", st_buffer_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_const_output_buffer", st_Buffer_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_has_beam_monitors", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_num_beam_monitors", st_buffer_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_beam_monitor_indices_begin", st_buffer_size_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_beam_monitor_indices_end", st_buffer_size_p, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_beam_monitor_index", st_buffer_size_t, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_has_elem_by_elem_config", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_elem_by_elem_config", ct.POINTER(st_ElemByElemConfig), [st_TrackJobNew_p])
_fn("st_TrackJobNew_is_elem_by_elem_config_rolling", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_default_elem_by_elem_config_rolling_flag", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_set_default_elem_by_elem_config_rolling_flag", None, [st_TrackJobNew_p, ct.c_bool])
_fn("st_TrackJobNew_get_elem_by_elem_config_order", ct.c_int64, [st_TrackJobNew_p])
_fn("st_TrackJobNew_get_default_elem_by_elem_config_order", ct.c_int64, [st_TrackJobNew_p])
_fn("st_TrackJobNew_set_default_elem_by_elem_config_order", None, [st_TrackJobNew_p, ct.c_int64])
_fn("st_TrackJobNew_add_assign_address_item", st_AssignAddressItem_p,
    [st_TrackJobNew_p, st_AssignAddressItem_p])
_fn("st_TrackJobNew_remove_assign_address_item", st_arch_status_t, [st_TrackJobNew_p, st_AssignAddressItem_p])
_fn("st_TrackJobNew_remove_assign_address_item_by_key_and_index", st_arch_status_t,
    [st_TrackJobNew_p, ct.POINTER(st_AssignAddressItemDestSrcPair), st_buffer_size_t])
_fn("st_TrackJobNew_has_assign_address_item", ct.c_bool, [st_TrackJobNew_p, st_AssignAddressItem_p])
_fn("st_TrackJobNew_has_assign_item_by_index", ct.c_bool,
    [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_has_assign_address_item_detailed", ct.c_bool,
    [st_TrackJobNew_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_index_of_assign_address_item_detailed", st_buffer_size_t,
    [st_TrackJobNew_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_index_of_assign_address_item", st_buffer_size_t,
    [st_TrackJobNew_p, st_AssignAddressItem_p])
_fn("st_TrackJobNew_has_assign_items", ct.c_bool, [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_num_assign_items", st_buffer_size_t,
    [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_ptr_assign_address_item", st_AssignAddressItem_p,
    [st_TrackJobNew_p, st_AssignAddressItem_p])
_fn("st_TrackJobNew_ptr_assign_address_item_detailed", st_AssignAddressItem_p,
    [st_TrackJobNew_p, st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t,
     st_object_type_id_t, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_ptr_assign_address_item_by_index", st_AssignAddressItem_p,
    [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_num_distinct_available_assign_address_items_dest_src_pairs", st_buffer_size_t,
    [st_TrackJobNew_p])
_fn("st_TrackJobNew_available_assign_address_items_dest_src_pairs", st_buffer_size_t,
    [st_TrackJobNew_p, st_buffer_size_t, ct.POINTER(st_AssignAddressItemDestSrcPair)])
_fn("st_TrackJobNew_buffer_by_buffer_id", st_Buffer_p, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_const_buffer_by_buffer_id", st_Buffer_p, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_is_buffer_by_buffer_id", ct.c_bool, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_is_raw_memory_by_buffer_id", ct.c_bool, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_assign_items_begin", ct.c_void_p, [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_assign_items_end", ct.c_void_p, [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_assign_item_dest_src_begin", ct.POINTER(st_AssignAddressItemDestSrcPair),
    [st_TrackJobNew_p])
_fn("st_TrackJobNew_assign_item_dest_src_end", ct.POINTER(st_AssignAddressItemDestSrcPair),
    [st_TrackJobNew_p])
_fn("st_TrackJobNew_commit_address_assignments", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_assign_all_addresses", st_arch_status_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_assign_addresses", st_arch_status_t,
    [st_TrackJobNew_p, st_buffer_size_t, st_buffer_size_t])
_fn("st_TrackJobNew_stored_buffers_capacity", st_arch_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_reserve_stored_buffers_capacity", st_arch_status_t, [st_TrackJobNew_p, st_arch_size_t])
_fn("st_TrackJobNew_has_stored_buffers", ct.c_bool, [st_TrackJobNew_p])
_fn("st_TrackJobNew_num_stored_buffers", st_arch_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_min_stored_buffer_id", st_arch_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_max_stored_buffer_id", st_arch_size_t, [st_TrackJobNew_p])
_fn("st_TrackJobNew_create_stored_buffer", st_arch_size_t, [st_TrackJobNew_p, st_buffer_size_t])
_fn("st_TrackJobNew_add_stored_buffer", st_arch_size_t, [st_TrackJobNew_p, st_Buffer_p, ct.c_bool, ct.c_bool])
_fn("st_TrackJobNew_owns_stored_buffer", ct.c_bool, [st_TrackJobNew_p, st_arch_size_t])
_fn("st_TrackJobNew_remove_stored_buffer", st_arch_status_t, [st_TrackJobNew_p, st_arch_size_t])
_fn("st_TrackJobNew_stored_buffer", st_Buffer_p, [st_TrackJobNew_p, st_arch_size_t])
_fn("st_TrackJobNew_const_stored_buffer", st_Buffer_p, [st_TrackJobNew_p, st_arch_size_t])
_fn("st_TrackJobNew_push_stored_buffer", st_arch_status_t, [st_TrackJobNew_p, st_arch_size_t])
_fn("st_TrackJobNew_collect_stored_buffer", st_arch_status_t, [st_TrackJobNew_p, st_arch_size_t])
_fn("st_CudaTrackJob_get_available_node_ids_list", ct.c_uint64, [ct.c_uint64, ct.POINTER(st_NodeId)])
_fn("st_CudaTrackJob_new_from_config_str", st_TrackJobNew_p, [ct.c_char_p])
_fn("st_CudaTrackJob_get_ptr_controller", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_const_controller", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_default_track_threads_per_block", st_arch_size_t, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_has_particles_arg", ct.c_bool, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_const_particles_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_has_beam_elements_arg", ct.c_bool, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_beam_elements_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_const_beam_elements_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_has_output_arg", ct.c_bool, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_output_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_const_output_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_has_elem_by_elem_config_arg", ct.c_bool, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_elem_by_elem_config_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_const_elem_by_elem_config_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_has_debug_register_arg", ct.c_bool, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_debug_register_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_const_debug_register_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_has_particles_addr_arg", ct.c_bool, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_particles_addr_arg", ct.c_void_p, [st_TrackJobNew_p])
_fn("st_CudaTrackJob_get_ptr_const_particles_addr_arg", ct.c_void_p, [st_TrackJobNew_p])

_fn("st_Hip_num_devices", ct.c_int, [])
_fn("st_Hip_last_error_str", ct.c_char_p, [])


def string_to_encoded_ctypes_str(s):
    return None if s is None else s.encode("utf-8")


def last_error():
    msg = st_Hip_last_error_str()
    return msg.decode() if msg else ""

/* sixtracklib_amd/track_job_new.h -- the new-style track-job C API and
 * the CUDA track job's constructors (drop-in boundary).
 *
 * Replaces sixtracklib/common/track/track_job_base.h:20-560 (the
 * NS(TrackJobNew_*) functions that the reference's GPU job, CudaTrackJob,
 * is driven through) and sixtracklib/cuda/track_job.h:20-170
 * (NS(CudaTrackJob_*)).  Both act on the job of track_job.h: a
 * NS(TrackJobBaseNew) or NS(CudaTrackJob) pointer is a NS(TrackJobBase)
 * pointer, and every function here has the meaning of the legacy function
 * of the same name, with the return conventions of the new API:
 *   - collect / collect_detailed return the flags collected, push_* the
 *     flags pushed;
 *   - collect_<part>, reset*, select_particle_set and assign_output_buffer
 *     return NS(arch_status_t) (0 success, -1 failure);
 *   - particle addresses are the device addresses of the job's copy of
 *     each particle set (legacy NS(TrackJob_fetch_particle_addresses)).
 * Architecture strings: "hip" (and its aliases, see track_job.h) and
 * "cuda", which the reference uses for its GPU job and which here names
 * the same HIP job; arch id ST_ARCHITECTURE_HIP.  The reference's
 * controller and argument objects (cuda/controller.h, cuda/argument.h)
 * have no counterpart: uses_controller / uses_arguments are false, the
 * CudaTrackJob argument getters return NULL and their has_* false.
 * Debug mode is accepted and changes nothing: map failures are reported
 * through the track status in either mode.
 */
#ifndef SIXTRACKLIB_AMD_TRACK_JOB_NEW_H__
#define SIXTRACKLIB_AMD_TRACK_JOB_NEW_H__

#include "sixtracklib_amd/track_job.h"

#if defined( __cplusplus )
extern "C" {
#endif

typedef NS(TrackJobBase) NS(TrackJobBaseNew);
typedef NS(TrackJobBase) NS(CudaTrackJob);
typedef uint64_t NS(arch_id_t);
typedef NS(arch_size_t) NS(ctrl_size_t);
typedef uint32_t NS(node_index_t);

/* control/definitions.h:61-90 (CPU 1, OpenCL 2, CUDA 3); HIP: this library */
#define ST_ARCHITECTURE_ILLEGAL ( ( NS(arch_id_t) )0x000003FFu )
#define ST_ARCHITECTURE_NONE    ( ( NS(arch_id_t) )0x00000000u )
#define ST_ARCHITECTURE_HIP     ( ( NS(arch_id_t) )0x00000004u )

/* common/track/track_job_base.h:20-560 */
ST_API NS(TrackJobBaseNew)* NS(TrackJobNew_create)( char const* arch_str, char const* config_str );
ST_API NS(TrackJobBaseNew)* NS(TrackJobNew_new)( char const* arch, NS(Buffer)* particles_buffer,
    NS(Buffer)* beam_elem_buffer, char const* config_str );
ST_API NS(TrackJobBaseNew)* NS(TrackJobNew_new_with_output)( char const* arch, NS(Buffer)* particles_buffer,
    NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem,
    char const* config_str );
ST_API NS(TrackJobBaseNew)* NS(TrackJobNew_new_detailed)( char const* arch, NS(Buffer)* particles_buffer,
    NS(buffer_size_t) num_particle_sets, NS(buffer_size_t) const* particle_set_indices_begin,
    NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem,
    char const* config_str );
ST_API void NS(TrackJobNew_delete)( NS(TrackJobBaseNew)* job );
ST_API NS(track_status_t) NS(TrackJobNew_track_until)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) until_turn );
ST_API NS(track_status_t) NS(TrackJobNew_track_elem_by_elem)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) until_turn );
ST_API NS(track_status_t) NS(TrackJobNew_track_line)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) beam_elem_begin_index, NS(buffer_size_t) beam_elem_end_index, bool finish_turn );
ST_API NS(track_job_collect_flag_t) NS(TrackJobNew_collect)( NS(TrackJobBaseNew)* job );
ST_API NS(track_job_collect_flag_t) NS(TrackJobNew_collect_detailed)( NS(TrackJobBaseNew)* job,
    NS(track_job_collect_flag_t) flags );
ST_API NS(arch_status_t) NS(TrackJobNew_collect_particles)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_collect_beam_elements)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_collect_output)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_collect_debug_flag)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_collect_particles_addresses)( NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_enable_collect_particles)( NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_disable_collect_particles)( NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_is_collecting_particles)( const NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_enable_collect_beam_elements)( NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_disable_collect_beam_elements)( NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_is_collecting_beam_elements)( const NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_enable_collect_output)( NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_disable_collect_output)( NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_is_collecting_output)( const NS(TrackJobBaseNew)* job );
ST_API NS(track_job_collect_flag_t) NS(TrackJobNew_get_collect_flags)( const NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_set_collect_flags)( NS(TrackJobBaseNew)* job, NS(track_job_collect_flag_t) flag );
ST_API bool NS(TrackJobNew_requires_collecting)( const NS(TrackJobBaseNew)* job );
ST_API NS(track_job_push_flag_t) NS(TrackJobNew_push)( NS(TrackJobBaseNew)* job,
    NS(track_job_push_flag_t) flag );
ST_API NS(track_job_push_flag_t) NS(TrackJobNew_push_particles)( NS(TrackJobBaseNew)* job );
ST_API NS(track_job_push_flag_t) NS(TrackJobNew_push_beam_elements)( NS(TrackJobBaseNew)* job );
ST_API NS(track_job_push_flag_t) NS(TrackJobNew_push_output)( NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_can_fetch_particle_addresses)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_has_particle_addresses)( const NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_fetch_particle_addresses)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_clear_particle_addresses)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) particle_set_index );
ST_API NS(arch_status_t) NS(TrackJobNew_clear_all_particle_addresses)( NS(TrackJobBaseNew)* job );
ST_API NS(ParticlesAddr) const* NS(TrackJobNew_get_particle_addresses)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) particle_set_index );
ST_API NS(Buffer) const* NS(TrackJobNew_get_ptr_particle_addresses_buffer)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_is_in_debug_mode)( const NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_enable_debug_mode)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_disable_debug_mode)( NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_clear)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_reset)( NS(TrackJobBaseNew)* job, NS(Buffer)* particles_buffer,
    NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer );
ST_API NS(arch_status_t) NS(TrackJobNew_reset_particle_set)( NS(TrackJobBaseNew)* job,
    NS(Buffer)* particles_buffer, NS(buffer_size_t) particle_set_index, NS(Buffer)* beam_elem_buffer,
    NS(Buffer)* output_buffer );
ST_API NS(arch_status_t) NS(TrackJobNew_reset_with_output)( NS(TrackJobBaseNew)* job,
    NS(Buffer)* particles_buffer, NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer,
    NS(buffer_size_t) until_turn_elem_by_elem );
ST_API NS(arch_status_t) NS(TrackJobNew_reset_detailed)( NS(TrackJobBaseNew)* job,
    NS(Buffer)* particles_buffer, NS(buffer_size_t) num_particle_sets,
    NS(buffer_size_t) const* particle_set_indices_begin, NS(Buffer)* beam_elem_buffer,
    NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem );
ST_API NS(arch_status_t) NS(TrackJobNew_select_particle_set)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) particle_set_index );
ST_API NS(arch_status_t) NS(TrackJobNew_assign_output_buffer)( NS(TrackJobBaseNew)* job,
    NS(Buffer)* ptr_output_buffer );
ST_API NS(arch_id_t) NS(TrackJobNew_get_arch_id)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_has_arch_string)( const NS(TrackJobBaseNew)* job );
ST_API char const* NS(TrackJobNew_get_arch_string)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_has_config_str)( const NS(TrackJobBaseNew)* job );
ST_API char const* NS(TrackJobNew_get_config_str)( const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_num_particle_sets)( const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) const* NS(TrackJobNew_get_particle_set_indices_begin)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) const* NS(TrackJobNew_get_particle_set_indices_end)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_particle_set_index)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) n );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_total_num_of_particles)( const NS(TrackJobBaseNew)* job );
ST_API NS(particle_index_t) NS(TrackJobNew_get_min_particle_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(particle_index_t) NS(TrackJobNew_get_max_particle_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(particle_index_t) NS(TrackJobNew_get_min_element_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(particle_index_t) NS(TrackJobNew_get_max_element_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(particle_index_t) NS(TrackJobNew_get_min_initial_turn_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(particle_index_t) NS(TrackJobNew_get_max_initial_turn_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(Buffer)* NS(TrackJobNew_get_particles_buffer)( NS(TrackJobBaseNew)* job );
ST_API NS(Buffer) const* NS(TrackJobNew_get_const_particles_buffer)( const NS(TrackJobBaseNew)* job );
ST_API NS(Buffer)* NS(TrackJobNew_get_beam_elements_buffer)( NS(TrackJobBaseNew)* job );
ST_API NS(Buffer) const* NS(TrackJobNew_get_const_beam_elements_buffer)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_has_output_buffer)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_owns_output_buffer)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_has_elem_by_elem_output)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_has_beam_monitor_output)( const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_beam_monitor_output_buffer_offset)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_elem_by_elem_output_buffer_offset)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_num_elem_by_elem_turns)( const NS(TrackJobBaseNew)* job );
ST_API NS(Buffer)* NS(TrackJobNew_get_output_buffer)( NS(TrackJobBaseNew)* job );
ST_API NS(Buffer) const* NS(TrackJobNew_get_const_output_buffer)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_has_beam_monitors)( const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_num_beam_monitors)( const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) const* NS(TrackJobNew_get_beam_monitor_indices_begin)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) const* NS(TrackJobNew_get_beam_monitor_indices_end)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_get_beam_monitor_index)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) n );
ST_API bool NS(TrackJobNew_has_elem_by_elem_config)( const NS(TrackJobBaseNew)* job );
ST_API NS(ElemByElemConfig) const* NS(TrackJobNew_get_elem_by_elem_config)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_is_elem_by_elem_config_rolling)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_get_default_elem_by_elem_config_rolling_flag)( const NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_set_default_elem_by_elem_config_rolling_flag)( NS(TrackJobBaseNew)* job,
    bool is_rolling_flag );
ST_API NS(elem_by_elem_order_t) NS(TrackJobNew_get_elem_by_elem_config_order)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(elem_by_elem_order_t) NS(TrackJobNew_get_default_elem_by_elem_config_order)(
    const NS(TrackJobBaseNew)* job );
ST_API void NS(TrackJobNew_set_default_elem_by_elem_config_order)( NS(TrackJobBaseNew)* job,
    NS(elem_by_elem_order_t) order );
ST_API bool NS(TrackJobNew_uses_controller)( const NS(TrackJobBaseNew)* job );
ST_API bool NS(TrackJobNew_uses_arguments)( const NS(TrackJobBaseNew)* job );
ST_API NS(AssignAddressItem)* NS(TrackJobNew_add_assign_address_item)( NS(TrackJobBaseNew)* job,
    const NS(AssignAddressItem)* assign_item_to_add );
ST_API NS(AssignAddressItem)* NS(TrackJobNew_add_assign_address_item_detailed)( NS(TrackJobBaseNew)* job,
    NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) dest_elem_index,
    NS(buffer_size_t) dest_pointer_offset, NS(object_type_id_t) src_type_id,
    NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset );
ST_API NS(arch_status_t) NS(TrackJobNew_remove_assign_address_item)( NS(TrackJobBaseNew)* job,
    NS(AssignAddressItem)* item_to_remove );
ST_API NS(arch_status_t) NS(TrackJobNew_remove_assign_address_item_by_key_and_index)(
    NS(TrackJobBaseNew)* job, const NS(TrackJobDestSrcBufferIds)* key,
    NS(buffer_size_t) index_of_item_to_remove );
ST_API bool NS(TrackJobNew_has_assign_address_item)( const NS(TrackJobBaseNew)* job,
    const NS(AssignAddressItem)* assign_item_to_add );
ST_API bool NS(TrackJobNew_has_assign_item_by_index)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) assign_item_index );
ST_API bool NS(TrackJobNew_has_assign_address_item_detailed)( const NS(TrackJobBaseNew)* job,
    NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) dest_elem_index,
    NS(buffer_size_t) dest_pointer_offset, NS(object_type_id_t) src_type_id,
    NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset );
ST_API NS(buffer_size_t) NS(TrackJobNew_index_of_assign_address_item_detailed)(
    const NS(TrackJobBaseNew)* job, NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index,
    NS(buffer_size_t) src_pointer_offset );
ST_API NS(buffer_size_t) NS(TrackJobNew_index_of_assign_address_item)( const NS(TrackJobBaseNew)* job,
    const NS(AssignAddressItem)* assign_item_to_add );
ST_API bool NS(TrackJobNew_has_assign_items)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );
ST_API NS(buffer_size_t) NS(TrackJobNew_num_assign_items)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );
ST_API NS(buffer_size_t) NS(TrackJobNew_total_num_assign_items)( const NS(TrackJobBaseNew)* job );
ST_API NS(AssignAddressItem) const* NS(TrackJobNew_ptr_assign_address_item)( const NS(TrackJobBaseNew)* job,
    const NS(AssignAddressItem)* assign_address_item );
ST_API NS(AssignAddressItem) const* NS(TrackJobNew_ptr_assign_address_item_detailed)(
    const NS(TrackJobBaseNew)* job, NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index,
    NS(buffer_size_t) src_pointer_offset );
ST_API NS(AssignAddressItem) const* NS(TrackJobNew_ptr_assign_address_item_by_index)(
    const NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id,
    NS(buffer_size_t) assign_address_item_index );
ST_API NS(buffer_size_t) NS(TrackJobNew_num_distinct_available_assign_address_items_dest_src_pairs)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(buffer_size_t) NS(TrackJobNew_available_assign_address_items_dest_src_pairs)(
    const NS(TrackJobBaseNew)* job, NS(buffer_size_t) max_num_pairs,
    NS(TrackJobDestSrcBufferIds)* pairs_begin );
ST_API NS(Buffer)* NS(TrackJobNew_buffer_by_buffer_id)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) buffer_id );
ST_API NS(Buffer) const* NS(TrackJobNew_const_buffer_by_buffer_id)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) buffer_id );
ST_API bool NS(TrackJobNew_is_buffer_by_buffer_id)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) buffer_id );
ST_API bool NS(TrackJobNew_is_raw_memory_by_buffer_id)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) buffer_id );
ST_API NS(Object) const* NS(TrackJobNew_assign_items_begin)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );
ST_API NS(Object) const* NS(TrackJobNew_assign_items_end)( const NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );
ST_API NS(TrackJobDestSrcBufferIds) const* NS(TrackJobNew_assign_item_dest_src_begin)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(TrackJobDestSrcBufferIds) const* NS(TrackJobNew_assign_item_dest_src_end)(
    const NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_commit_address_assignments)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_assign_all_addresses)( NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_assign_addresses)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );
ST_API NS(arch_size_t) NS(TrackJobNew_stored_buffers_capacity)( const NS(TrackJobBaseNew)* job );
ST_API NS(arch_status_t) NS(TrackJobNew_reserve_stored_buffers_capacity)( NS(TrackJobBaseNew)* job,
    NS(arch_size_t) capacity );
ST_API bool NS(TrackJobNew_has_stored_buffers)( const NS(TrackJobBaseNew)* job );
ST_API NS(arch_size_t) NS(TrackJobNew_num_stored_buffers)( const NS(TrackJobBaseNew)* job );
ST_API NS(arch_size_t) NS(TrackJobNew_min_stored_buffer_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(arch_size_t) NS(TrackJobNew_max_stored_buffer_id)( const NS(TrackJobBaseNew)* job );
ST_API NS(arch_size_t) NS(TrackJobNew_create_stored_buffer)( NS(TrackJobBaseNew)* job,
    NS(buffer_size_t) buffer_capacity );
ST_API NS(arch_size_t) NS(TrackJobNew_add_stored_buffer)( NS(TrackJobBaseNew)* job, NS(Buffer)* buffer,
    bool take_ownership, bool delete_ptr_after_move );
ST_API bool NS(TrackJobNew_owns_stored_buffer)( const NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_id );
ST_API NS(arch_status_t) NS(TrackJobNew_remove_stored_buffer)( NS(TrackJobBaseNew)* job,
    NS(arch_size_t) buffer_index );
ST_API NS(Buffer)* NS(TrackJobNew_stored_buffer)( NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_id );
ST_API NS(Buffer) const* NS(TrackJobNew_const_stored_buffer)( const NS(TrackJobBaseNew)* job,
    NS(arch_size_t) buffer_id );
ST_API NS(arch_status_t) NS(TrackJobNew_push_stored_buffer)( NS(TrackJobBaseNew)* job,
    NS(arch_size_t) buffer_id );
ST_API NS(arch_status_t) NS(TrackJobNew_collect_stored_buffer)( NS(TrackJobBaseNew)* job,
    NS(arch_size_t) buffer_id );

/* cuda/track_job.h:20-170.  A node is a HIP device: its index is the
 * device number, its id string "0.N" (platform 0, device N; "N" alone is
 * accepted too).  NS(NodeId) is opaque in the reference's C API; here it
 * is the pair below. */
typedef struct NS(NodeId)
{
    int64_t platform_id;
    int64_t device_id;
    NS(node_index_t) node_index;
}
NS(NodeId);

ST_API NS(ctrl_size_t) NS(CudaTrackJob_get_num_available_nodes)( void );
ST_API NS(ctrl_size_t) NS(CudaTrackJob_get_available_node_ids_list)( NS(ctrl_size_t) max_num_node_ids,
    NS(NodeId)* node_ids_begin );
ST_API NS(ctrl_size_t) NS(CudaTrackJob_get_available_node_indices_list)(
    NS(ctrl_size_t) max_num_node_indices, NS(node_index_t)* node_indices_begin );
ST_API NS(CudaTrackJob)* NS(CudaTrackJob_create)( void );
ST_API NS(CudaTrackJob)* NS(CudaTrackJob_new_from_config_str)( char const* config_str );
ST_API NS(CudaTrackJob)* NS(CudaTrackJob_new)( char const* node_id_str,
    NS(Buffer)* particles_buffer, NS(Buffer)* beam_elements_buffer );
ST_API NS(CudaTrackJob)* NS(CudaTrackJob_new_with_output)( char const* node_id_str,
    NS(Buffer)* particles_buffer, NS(Buffer)* beam_elements_buffer, NS(Buffer)* output_buffer,
    NS(buffer_size_t) until_turn_elem_by_elem );
ST_API NS(CudaTrackJob)* NS(CudaTrackJob_new_detailed)( char const* node_id_str,
    NS(Buffer)* particles_buffer, NS(buffer_size_t) num_psets, NS(buffer_size_t) const* pset_indices_begin,
    NS(Buffer)* beam_elements_buffer, NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem,
    char const* config_str );
ST_API bool NS(CudaTrackJob_has_controller)( const NS(CudaTrackJob)* track_job );
ST_API void* NS(CudaTrackJob_get_ptr_controller)( NS(CudaTrackJob)* track_job );
ST_API void const* NS(CudaTrackJob_get_ptr_const_controller)( const NS(CudaTrackJob)* track_job );
#define ST_CUDA_TRACK_JOB_ARG_GETTERS( what ) \
    ST_API bool NS(CudaTrackJob_has_##what##_arg)( const NS(CudaTrackJob)* track_job ); \
    ST_API void* NS(CudaTrackJob_get_ptr_##what##_arg)( NS(CudaTrackJob)* track_job ); \
    ST_API void const* NS(CudaTrackJob_get_ptr_const_##what##_arg)( const NS(CudaTrackJob)* track_job );
ST_CUDA_TRACK_JOB_ARG_GETTERS( particles )
ST_CUDA_TRACK_JOB_ARG_GETTERS( beam_elements )
ST_CUDA_TRACK_JOB_ARG_GETTERS( output )
ST_CUDA_TRACK_JOB_ARG_GETTERS( elem_by_elem_config )
ST_CUDA_TRACK_JOB_ARG_GETTERS( debug_register )
ST_CUDA_TRACK_JOB_ARG_GETTERS( particles_addr )
#undef ST_CUDA_TRACK_JOB_ARG_GETTERS
/* threads per block of the job's launches (config token block=) */
ST_API NS(arch_size_t) NS(CudaTrackJob_default_threads_per_block)( const NS(CudaTrackJob)* track_job );
ST_API NS(arch_size_t) NS(CudaTrackJob_default_track_threads_per_block)( const NS(CudaTrackJob)* track_job );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_TRACK_JOB_NEW_H__ */

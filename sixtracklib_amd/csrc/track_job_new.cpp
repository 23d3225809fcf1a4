// NS(TrackJobNew_*) and NS(CudaTrackJob_*) (include/sixtracklib_amd/
// track_job_new.h): the reference's new-style job API
// (common/track/track_job_base.h:20-560, implemented in
// common/track/track_job_base_c99.cpp) and its CUDA job's constructors
// (cuda/track_job.h:20-170) as forwards to the one job of track_job.cpp.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <string>

#include "sixtracklib_amd/track_job_new.h"

// the debug-mode flag kept in the job (track_job.cpp)
extern "C" bool st_internal_job_debug_mode( const NS(TrackJobBase)* j );
extern "C" void st_internal_job_set_debug_mode( NS(TrackJobBase)* j, bool on );

namespace
{
inline NS(TrackJobBase)* J( NS(TrackJobBaseNew)* j ) { return j; }
inline const NS(TrackJobBase)* J( const NS(TrackJobBaseNew)* j ) { return j; }

// the reference's GPU arch string "cuda" names this library's HIP job
char const* hip_arch( char const* arch )
{
    if( arch == nullptr ) return nullptr;
    std::string a;
    for( char const* c = arch; *c; ++c ) a.push_back( ( char )std::tolower( ( unsigned char )*c ) );
    return ( a == "cuda" || a.rfind( "cuda:", 0 ) == 0 ) ? "hip" : arch;
}

NS(TrackJobBaseNew)* job_new_from( char const* arch, char const* config )
{
    return NS(TrackJob_create)( hip_arch( arch ), config );
}

// node id "P.D" or "D" -> "device=D" appended to the configuration
std::string node_config( char const* node_id, char const* config )
{
    std::string c = config ? config : "";
    if( node_id != nullptr && *node_id )
    {
        char const* dot = std::strchr( node_id, '.' );
        long const dev = std::strtol( dot ? dot + 1 : node_id, nullptr, 10 );
        c += ( c.empty() ? "" : " " ) + std::string( "device=" ) + std::to_string( dev );
    }
    return c;
}
}  // namespace

extern "C" {

NS(TrackJobBaseNew)* NS(TrackJobNew_create)( char const* arch_str, char const* config_str )
{
    return job_new_from( arch_str, config_str );
}

NS(TrackJobBaseNew)* NS(TrackJobNew_new)( char const* arch, NS(Buffer)* particles_buffer, NS(Buffer)* beam_elem_buffer, char const* config_str )
{
    return reinterpret_cast< NS(TrackJobBaseNew)* >( NS(TrackJob_new)( hip_arch( arch ), particles_buffer, beam_elem_buffer, config_str ) );
}

NS(TrackJobBaseNew)* NS(TrackJobNew_new_with_output)( char const* arch, NS(Buffer)* particles_buffer, NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem, char const* config_str )
{
    return reinterpret_cast< NS(TrackJobBaseNew)* >( NS(TrackJob_new_with_output)( hip_arch( arch ), particles_buffer, beam_elem_buffer, output_buffer, until_turn_elem_by_elem, config_str ) );
}

NS(TrackJobBaseNew)* NS(TrackJobNew_new_detailed)( char const* arch, NS(Buffer)* particles_buffer, NS(buffer_size_t) num_particle_sets, NS(buffer_size_t) const* particle_set_indices_begin, NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem, char const* config_str )
{
    return reinterpret_cast< NS(TrackJobBaseNew)* >( NS(TrackJob_new_detailed)( hip_arch( arch ), particles_buffer, num_particle_sets, particle_set_indices_begin, beam_elem_buffer, output_buffer, until_turn_elem_by_elem, config_str ) );
}

void NS(TrackJobNew_delete)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_delete)( J( job ) );
}

NS(track_status_t) NS(TrackJobNew_track_until)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) until_turn )
{
    return NS(TrackJob_track_until)( J( job ), until_turn );
}

NS(track_status_t) NS(TrackJobNew_track_elem_by_elem)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) until_turn )
{
    return NS(TrackJob_track_elem_by_elem)( J( job ), until_turn );
}

NS(track_status_t) NS(TrackJobNew_track_line)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) beam_elem_begin_index, NS(buffer_size_t) beam_elem_end_index, bool finish_turn )
{
    return NS(TrackJob_track_line)( J( job ), beam_elem_begin_index, beam_elem_end_index, finish_turn );
}

NS(track_job_collect_flag_t) NS(TrackJobNew_collect)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return 0;
    NS(track_job_collect_flag_t) const f = NS(TrackJob_get_collect_flags)( J( job ) );
    NS(TrackJob_collect)( J( job ) );
    return f;
}

NS(track_job_collect_flag_t) NS(TrackJobNew_collect_detailed)( NS(TrackJobBaseNew)* job, NS(track_job_collect_flag_t) flags )
{
    if( job == nullptr ) return 0;
    NS(TrackJob_collect_detailed)( J( job ), flags );
    return flags;
}

NS(arch_status_t) NS(TrackJobNew_collect_particles)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    NS(TrackJob_collect_particles)( J( job ) );
    return ST_ARCH_STATUS_SUCCESS;
}

NS(arch_status_t) NS(TrackJobNew_collect_beam_elements)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    NS(TrackJob_collect_beam_elements)( J( job ) );
    return ST_ARCH_STATUS_SUCCESS;
}

NS(arch_status_t) NS(TrackJobNew_collect_output)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    NS(TrackJob_collect_output)( J( job ) );
    return ST_ARCH_STATUS_SUCCESS;
}

NS(arch_status_t) NS(TrackJobNew_collect_debug_flag)( NS(TrackJobBaseNew)* job )
{
    return job ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJobNew_collect_particles_addresses)( NS(TrackJobBaseNew)* job )
{
    return job ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

void NS(TrackJobNew_enable_collect_particles)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_enable_collect_particles)( J( job ) );
}

void NS(TrackJobNew_disable_collect_particles)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_disable_collect_particles)( J( job ) );
}

bool NS(TrackJobNew_is_collecting_particles)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_is_collecting_particles)( J( job ) );
}

void NS(TrackJobNew_enable_collect_beam_elements)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_enable_collect_beam_elements)( J( job ) );
}

void NS(TrackJobNew_disable_collect_beam_elements)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_disable_collect_beam_elements)( J( job ) );
}

bool NS(TrackJobNew_is_collecting_beam_elements)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_is_collecting_beam_elements)( J( job ) );
}

void NS(TrackJobNew_enable_collect_output)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_enable_collect_output)( J( job ) );
}

void NS(TrackJobNew_disable_collect_output)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_disable_collect_output)( J( job ) );
}

bool NS(TrackJobNew_is_collecting_output)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_is_collecting_output)( J( job ) );
}

NS(track_job_collect_flag_t) NS(TrackJobNew_get_collect_flags)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_collect_flags)( J( job ) );
}

void NS(TrackJobNew_set_collect_flags)( NS(TrackJobBaseNew)* job, NS(track_job_collect_flag_t) flag )
{
    NS(TrackJob_set_collect_flags)( J( job ), flag );
}

bool NS(TrackJobNew_requires_collecting)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_requires_collecting)( J( job ) );
}

NS(track_job_push_flag_t) NS(TrackJobNew_push)( NS(TrackJobBaseNew)* job, NS(track_job_push_flag_t) flag )
{
    if( job == nullptr ) return 0;
    NS(TrackJob_push)( J( job ), flag );
    return flag;
}

NS(track_job_push_flag_t) NS(TrackJobNew_push_particles)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return 0;
    NS(TrackJob_push_particles)( J( job ) );
    return ST_TRACK_JOB_IO_PARTICLES;
}

NS(track_job_push_flag_t) NS(TrackJobNew_push_beam_elements)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return 0;
    NS(TrackJob_push_beam_elements)( J( job ) );
    return ST_TRACK_JOB_IO_BEAM_ELEMENTS;
}

NS(track_job_push_flag_t) NS(TrackJobNew_push_output)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return 0;
    NS(TrackJob_push_output)( J( job ) );
    return ST_TRACK_JOB_IO_OUTPUT;
}

bool NS(TrackJobNew_can_fetch_particle_addresses)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_can_fetch_particles_addr)( J( job ) );
}

bool NS(TrackJobNew_has_particle_addresses)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_particles_addr)( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_fetch_particle_addresses)( NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_fetch_particle_addresses)( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_clear_particle_addresses)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) particle_set_index )
{
    return NS(TrackJob_clear_particle_addresses)( J( job ), particle_set_index );
}

NS(arch_status_t) NS(TrackJobNew_clear_all_particle_addresses)( NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_clear_all_particle_addresses)( J( job ) );
}

NS(ParticlesAddr) const* NS(TrackJobNew_get_particle_addresses)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) particle_set_index )
{
    return NS(TrackJob_particle_addresses)( J( job ), particle_set_index );
}

NS(Buffer) const* NS(TrackJobNew_get_ptr_particle_addresses_buffer)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_const_particles_addr_buffer)( J( job ) );
}

bool NS(TrackJobNew_is_in_debug_mode)( const NS(TrackJobBaseNew)* job )
{
    return st_internal_job_debug_mode( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_enable_debug_mode)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    st_internal_job_set_debug_mode( J( job ), true );
    return ST_ARCH_STATUS_SUCCESS;
}

NS(arch_status_t) NS(TrackJobNew_disable_debug_mode)( NS(TrackJobBaseNew)* job )
{
    if( job == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    st_internal_job_set_debug_mode( J( job ), false );
    return ST_ARCH_STATUS_SUCCESS;
}

void NS(TrackJobNew_clear)( NS(TrackJobBaseNew)* job )
{
    NS(TrackJob_clear)( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_reset)( NS(TrackJobBaseNew)* job, NS(Buffer)* particles_buffer, NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer )
{
    return NS(TrackJob_reset)( J( job ), particles_buffer, beam_elem_buffer, output_buffer ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJobNew_reset_particle_set)( NS(TrackJobBaseNew)* job, NS(Buffer)* particles_buffer, NS(buffer_size_t) particle_set_index, NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer )
{
    return NS(TrackJob_reset_particle_set)( J( job ), particles_buffer, particle_set_index, beam_elem_buffer, output_buffer ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJobNew_reset_with_output)( NS(TrackJobBaseNew)* job, NS(Buffer)* particles_buffer, NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem )
{
    return NS(TrackJob_reset_with_output)( J( job ), particles_buffer, beam_elem_buffer, output_buffer, until_turn_elem_by_elem ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJobNew_reset_detailed)( NS(TrackJobBaseNew)* job, NS(Buffer)* particles_buffer, NS(buffer_size_t) num_particle_sets, NS(buffer_size_t) const* particle_set_indices_begin, NS(Buffer)* beam_elem_buffer, NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem )
{
    return NS(TrackJob_reset_detailed)( J( job ), particles_buffer, num_particle_sets, particle_set_indices_begin, beam_elem_buffer, output_buffer, until_turn_elem_by_elem ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJobNew_select_particle_set)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) particle_set_index )
{
    return NS(TrackJob_select_particle_set)( J( job ), particle_set_index ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJobNew_assign_output_buffer)( NS(TrackJobBaseNew)* job, NS(Buffer)* ptr_output_buffer )
{
    return NS(TrackJob_assign_output_buffer)( J( job ), ptr_output_buffer ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_id_t) NS(TrackJobNew_get_arch_id)( const NS(TrackJobBaseNew)* job )
{
    return job ? ST_ARCHITECTURE_HIP : ST_ARCHITECTURE_ILLEGAL;
}

bool NS(TrackJobNew_has_arch_string)( const NS(TrackJobBaseNew)* job )
{
    return job != nullptr;
}

char const* NS(TrackJobNew_get_arch_string)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_type_str)( J( job ) );
}

bool NS(TrackJobNew_has_config_str)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_config_str)( J( job ) );
}

char const* NS(TrackJobNew_get_config_str)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_config_str)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_get_num_particle_sets)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_num_particle_sets)( J( job ) );
}

NS(buffer_size_t) const* NS(TrackJobNew_get_particle_set_indices_begin)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_particle_set_indices_begin)( J( job ) );
}

NS(buffer_size_t) const* NS(TrackJobNew_get_particle_set_indices_end)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_particle_set_indices_end)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_get_particle_set_index)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) n )
{
    return NS(TrackJob_get_particle_set_index)( J( job ), n );
}

NS(buffer_size_t) NS(TrackJobNew_get_total_num_of_particles)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_total_num_particles_in_sets)( J( job ) );
}

NS(particle_index_t) NS(TrackJobNew_get_min_particle_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_min_particle_id)( J( job ) );
}

NS(particle_index_t) NS(TrackJobNew_get_max_particle_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_max_particle_id)( J( job ) );
}

NS(particle_index_t) NS(TrackJobNew_get_min_element_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_min_element_id)( J( job ) );
}

NS(particle_index_t) NS(TrackJobNew_get_max_element_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_max_element_id)( J( job ) );
}

NS(particle_index_t) NS(TrackJobNew_get_min_initial_turn_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_min_initial_turn_id)( J( job ) );
}

NS(particle_index_t) NS(TrackJobNew_get_max_initial_turn_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_max_initial_turn_id)( J( job ) );
}

NS(Buffer)* NS(TrackJobNew_get_particles_buffer)( NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_particles_buffer)( J( job ) );
}

NS(Buffer) const* NS(TrackJobNew_get_const_particles_buffer)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_const_particles_buffer)( J( job ) );
}

NS(Buffer)* NS(TrackJobNew_get_beam_elements_buffer)( NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_beam_elements_buffer)( J( job ) );
}

NS(Buffer) const* NS(TrackJobNew_get_const_beam_elements_buffer)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_const_beam_elements_buffer)( J( job ) );
}

bool NS(TrackJobNew_has_output_buffer)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_output_buffer)( J( job ) );
}

bool NS(TrackJobNew_owns_output_buffer)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_owns_output_buffer)( J( job ) );
}

bool NS(TrackJobNew_has_elem_by_elem_output)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_elem_by_elem_output)( J( job ) );
}

bool NS(TrackJobNew_has_beam_monitor_output)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_beam_monitor_output)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_get_beam_monitor_output_buffer_offset)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_beam_monitor_output_buffer_offset)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_get_elem_by_elem_output_buffer_offset)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_elem_by_elem_output_buffer_offset)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_get_num_elem_by_elem_turns)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_num_elem_by_elem_turns)( J( job ) );
}

NS(Buffer)* NS(TrackJobNew_get_output_buffer)( NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_output_buffer)( J( job ) );
}

NS(Buffer) const* NS(TrackJobNew_get_const_output_buffer)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_const_output_buffer)( J( job ) );
}

bool NS(TrackJobNew_has_beam_monitors)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_beam_monitors)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_get_num_beam_monitors)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_num_beam_monitors)( J( job ) );
}

NS(buffer_size_t) const* NS(TrackJobNew_get_beam_monitor_indices_begin)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_beam_monitor_indices_begin)( J( job ) );
}

NS(buffer_size_t) const* NS(TrackJobNew_get_beam_monitor_indices_end)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_beam_monitor_indices_end)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_get_beam_monitor_index)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) n )
{
    return NS(TrackJob_get_beam_monitor_index)( J( job ), n );
}

bool NS(TrackJobNew_has_elem_by_elem_config)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_elem_by_elem_config)( J( job ) );
}

NS(ElemByElemConfig) const* NS(TrackJobNew_get_elem_by_elem_config)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_elem_by_elem_config)( J( job ) );
}

bool NS(TrackJobNew_is_elem_by_elem_config_rolling)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_is_elem_by_elem_config_rolling)( J( job ) );
}

bool NS(TrackJobNew_get_default_elem_by_elem_config_rolling_flag)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_default_elem_by_elem_config_rolling_flag)( J( job ) );
}

void NS(TrackJobNew_set_default_elem_by_elem_config_rolling_flag)( NS(TrackJobBaseNew)* job, bool is_rolling_flag )
{
    NS(TrackJob_set_default_elem_by_elem_config_rolling_flag)( J( job ), is_rolling_flag );
}

NS(elem_by_elem_order_t) NS(TrackJobNew_get_elem_by_elem_config_order)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_elem_by_elem_config_order)( J( job ) );
}

NS(elem_by_elem_order_t) NS(TrackJobNew_get_default_elem_by_elem_config_order)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_get_default_elem_by_elem_config_order)( J( job ) );
}

void NS(TrackJobNew_set_default_elem_by_elem_config_order)( NS(TrackJobBaseNew)* job, NS(elem_by_elem_order_t) order )
{
    NS(TrackJob_set_default_elem_by_elem_config_order)( J( job ), order );
}

bool NS(TrackJobNew_uses_controller)( const NS(TrackJobBaseNew)* job )
{
    return false;
}

bool NS(TrackJobNew_uses_arguments)( const NS(TrackJobBaseNew)* job )
{
    return false;
}

NS(AssignAddressItem)* NS(TrackJobNew_add_assign_address_item)( NS(TrackJobBaseNew)* job, const NS(AssignAddressItem)* assign_item_to_add )
{
    return NS(TrackJob_add_assign_address_item)( J( job ), assign_item_to_add );
}

NS(AssignAddressItem)* NS(TrackJobNew_add_assign_address_item_detailed)( NS(TrackJobBaseNew)* job, NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset, NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset )
{
    return NS(TrackJob_add_assign_address_item_detailed)( J( job ), dest_type_id, dest_buffer_id, dest_elem_index, dest_pointer_offset, src_type_id, src_buffer_id, src_elem_index, src_pointer_offset );
}

NS(arch_status_t) NS(TrackJobNew_remove_assign_address_item)( NS(TrackJobBaseNew)* job, NS(AssignAddressItem)* item_to_remove )
{
    return NS(TrackJob_remove_assign_address_item)( J( job ), item_to_remove );
}

NS(arch_status_t) NS(TrackJobNew_remove_assign_address_item_by_key_and_index)( NS(TrackJobBaseNew)* job, const NS(TrackJobDestSrcBufferIds)* key, NS(buffer_size_t) index_of_item_to_remove )
{
    return NS(TrackJob_remove_assign_address_item_by_key_and_index)( J( job ), key, index_of_item_to_remove );
}

bool NS(TrackJobNew_has_assign_address_item)( const NS(TrackJobBaseNew)* job, const NS(AssignAddressItem)* assign_item_to_add )
{
    return NS(TrackJob_has_assign_address_item)( J( job ), assign_item_to_add );
}

bool NS(TrackJobNew_has_assign_item_by_index)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) assign_item_index )
{
    return NS(TrackJob_has_assign_item_by_index)( J( job ), dest_buffer_id, src_buffer_id, assign_item_index );
}

bool NS(TrackJobNew_has_assign_address_item_detailed)( const NS(TrackJobBaseNew)* job, NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset, NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset )
{
    return NS(TrackJob_has_assign_address_item_detailed)( J( job ), dest_type_id, dest_buffer_id, dest_elem_index, dest_pointer_offset, src_type_id, src_buffer_id, src_elem_index, src_pointer_offset );
}

NS(buffer_size_t) NS(TrackJobNew_index_of_assign_address_item_detailed)( const NS(TrackJobBaseNew)* job, NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset, NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset )
{
    return NS(TrackJob_index_of_assign_address_item_detailed)( J( job ), dest_type_id, dest_buffer_id, dest_elem_index, dest_pointer_offset, src_type_id, src_buffer_id, src_elem_index, src_pointer_offset );
}

NS(buffer_size_t) NS(TrackJobNew_index_of_assign_address_item)( const NS(TrackJobBaseNew)* job, const NS(AssignAddressItem)* assign_item_to_add )
{
    return NS(TrackJob_index_of_assign_address_item)( J( job ), assign_item_to_add );
}

bool NS(TrackJobNew_has_assign_items)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id )
{
    return NS(TrackJob_has_assign_items)( J( job ), dest_buffer_id, src_buffer_id );
}

NS(buffer_size_t) NS(TrackJobNew_num_assign_items)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id )
{
    return NS(TrackJob_num_assign_items)( J( job ), dest_buffer_id, src_buffer_id );
}

NS(buffer_size_t) NS(TrackJobNew_total_num_assign_items)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_total_num_assign_items)( J( job ) );
}

NS(AssignAddressItem) const* NS(TrackJobNew_ptr_assign_address_item)( const NS(TrackJobBaseNew)* job, const NS(AssignAddressItem)* assign_address_item )
{
    return NS(TrackJob_ptr_assign_address_item)( J( job ), assign_address_item );
}

NS(AssignAddressItem) const* NS(TrackJobNew_ptr_assign_address_item_detailed)( const NS(TrackJobBaseNew)* job, NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset, NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset )
{
    return NS(TrackJob_ptr_assign_address_item_detailed)( J( job ), dest_type_id, dest_buffer_id, dest_elem_index, dest_pointer_offset, src_type_id, src_buffer_id, src_elem_index, src_pointer_offset );
}

NS(AssignAddressItem) const* NS(TrackJobNew_ptr_assign_address_item_by_index)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) assign_address_item_index )
{
    return NS(TrackJob_ptr_assign_address_item_by_index)( J( job ), dest_buffer_id, src_buffer_id, assign_address_item_index );
}

NS(buffer_size_t) NS(TrackJobNew_num_distinct_available_assign_address_items_dest_src_pairs)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_num_distinct_available_assign_address_items_dest_src_pairs)( J( job ) );
}

NS(buffer_size_t) NS(TrackJobNew_available_assign_address_items_dest_src_pairs)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) max_num_pairs, NS(TrackJobDestSrcBufferIds)* pairs_begin )
{
    return NS(TrackJob_available_assign_address_items_dest_src_pairs)( J( job ), max_num_pairs, pairs_begin );
}

NS(Buffer)* NS(TrackJobNew_buffer_by_buffer_id)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) buffer_id )
{
    return NS(TrackJob_buffer_by_buffer_id)( J( job ), buffer_id );
}

NS(Buffer) const* NS(TrackJobNew_const_buffer_by_buffer_id)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) buffer_id )
{
    return NS(TrackJob_const_buffer_by_buffer_id)( J( job ), buffer_id );
}

bool NS(TrackJobNew_is_buffer_by_buffer_id)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) buffer_id )
{
    return NS(TrackJob_is_buffer_by_buffer_id)( J( job ), buffer_id );
}

bool NS(TrackJobNew_is_raw_memory_by_buffer_id)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) buffer_id )
{
    return NS(TrackJob_is_raw_memory_by_buffer_id)( J( job ), buffer_id );
}

NS(Object) const* NS(TrackJobNew_assign_items_begin)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id )
{
    return NS(TrackJob_assign_items_begin)( J( job ), dest_buffer_id, src_buffer_id );
}

NS(Object) const* NS(TrackJobNew_assign_items_end)( const NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id )
{
    return NS(TrackJob_assign_items_end)( J( job ), dest_buffer_id, src_buffer_id );
}

NS(TrackJobDestSrcBufferIds) const* NS(TrackJobNew_assign_item_dest_src_begin)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_assign_item_dest_src_begin)( J( job ) );
}

NS(TrackJobDestSrcBufferIds) const* NS(TrackJobNew_assign_item_dest_src_end)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_assign_item_dest_src_end)( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_commit_address_assignments)( NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_commit_address_assignments)( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_assign_all_addresses)( NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_assign_all_addresses)( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_assign_addresses)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id )
{
    return NS(TrackJob_assign_addresses)( J( job ), dest_buffer_id, src_buffer_id );
}

NS(arch_size_t) NS(TrackJobNew_stored_buffers_capacity)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_stored_buffers_capacity)( J( job ) );
}

NS(arch_status_t) NS(TrackJobNew_reserve_stored_buffers_capacity)( NS(TrackJobBaseNew)* job, NS(arch_size_t) capacity )
{
    return NS(TrackJob_reserve_stored_buffers_capacity)( J( job ), capacity );
}

bool NS(TrackJobNew_has_stored_buffers)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_has_stored_buffers)( J( job ) );
}

NS(arch_size_t) NS(TrackJobNew_num_stored_buffers)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_num_stored_buffers)( J( job ) );
}

NS(arch_size_t) NS(TrackJobNew_min_stored_buffer_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_min_stored_buffer_id)( J( job ) );
}

NS(arch_size_t) NS(TrackJobNew_max_stored_buffer_id)( const NS(TrackJobBaseNew)* job )
{
    return NS(TrackJob_max_stored_buffer_id)( J( job ) );
}

NS(arch_size_t) NS(TrackJobNew_create_stored_buffer)( NS(TrackJobBaseNew)* job, NS(buffer_size_t) buffer_capacity )
{
    return NS(TrackJob_create_stored_buffer)( J( job ), buffer_capacity );
}

NS(arch_size_t) NS(TrackJobNew_add_stored_buffer)( NS(TrackJobBaseNew)* job, NS(Buffer)* buffer, bool take_ownership, bool delete_ptr_after_move )
{
    return NS(TrackJob_add_stored_buffer)( J( job ), buffer, take_ownership, delete_ptr_after_move );
}

bool NS(TrackJobNew_owns_stored_buffer)( const NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_id )
{
    return NS(TrackJob_owns_stored_buffer)( J( job ), buffer_id );
}

NS(arch_status_t) NS(TrackJobNew_remove_stored_buffer)( NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_index )
{
    return NS(TrackJob_remove_stored_buffer)( J( job ), buffer_index );
}

NS(Buffer)* NS(TrackJobNew_stored_buffer)( NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_id )
{
    return NS(TrackJob_stored_buffer)( J( job ), buffer_id );
}

NS(Buffer) const* NS(TrackJobNew_const_stored_buffer)( const NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_id )
{
    return NS(TrackJob_const_stored_buffer)( J( job ), buffer_id );
}

NS(arch_status_t) NS(TrackJobNew_push_stored_buffer)( NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_id )
{
    return NS(TrackJob_push_stored_buffer)( J( job ), buffer_id );
}

NS(arch_status_t) NS(TrackJobNew_collect_stored_buffer)( NS(TrackJobBaseNew)* job, NS(arch_size_t) buffer_id )
{
    return NS(TrackJob_collect_stored_buffer)( J( job ), buffer_id );
}

// -- cuda/track_job.h ---------------------------------------------------------
NS(ctrl_size_t) NS(CudaTrackJob_get_num_available_nodes)( void )
{
    return ( NS(ctrl_size_t) )NS(Hip_num_devices)();
}

NS(ctrl_size_t) NS(CudaTrackJob_get_available_node_ids_list)( NS(ctrl_size_t) max_num, NS(NodeId)* ids )
{
    NS(ctrl_size_t) const n = NS(CudaTrackJob_get_num_available_nodes)();
    NS(ctrl_size_t) k = 0;
    for( ; ids != nullptr && k < n && k < max_num; ++k )
    {
        ids[ k ].platform_id = 0;
        ids[ k ].device_id = ( int64_t )k;
        ids[ k ].node_index = ( NS(node_index_t) )k;
    }
    return k;
}

NS(ctrl_size_t) NS(CudaTrackJob_get_available_node_indices_list)( NS(ctrl_size_t) max_num,
    NS(node_index_t)* idx )
{
    NS(ctrl_size_t) const n = NS(CudaTrackJob_get_num_available_nodes)();
    NS(ctrl_size_t) k = 0;
    for( ; idx != nullptr && k < n && k < max_num; ++k ) idx[ k ] = ( NS(node_index_t) )k;
    return k;
}

NS(CudaTrackJob)* NS(CudaTrackJob_create)( void ) { return NS(TrackJob_create)( "hip", nullptr ); }

NS(CudaTrackJob)* NS(CudaTrackJob_new_from_config_str)( char const* config_str )
{
    return NS(TrackJob_create)( "hip", config_str );
}

NS(CudaTrackJob)* NS(CudaTrackJob_new)( char const* node_id_str, NS(Buffer)* pbuf, NS(Buffer)* ebuf )
{
    std::string const c = node_config( node_id_str, nullptr );
    return NS(TrackJob_new)( "hip", pbuf, ebuf, c.c_str() );
}

NS(CudaTrackJob)* NS(CudaTrackJob_new_with_output)( char const* node_id_str, NS(Buffer)* pbuf,
    NS(Buffer)* ebuf, NS(Buffer)* obuf, NS(buffer_size_t) until_turn_ebe )
{
    std::string const c = node_config( node_id_str, nullptr );
    return NS(TrackJob_new_with_output)( "hip", pbuf, ebuf, obuf, until_turn_ebe, c.c_str() );
}

NS(CudaTrackJob)* NS(CudaTrackJob_new_detailed)( char const* node_id_str, NS(Buffer)* pbuf,
    NS(buffer_size_t) num_psets, NS(buffer_size_t) const* pset_indices, NS(Buffer)* ebuf, NS(Buffer)* obuf,
    NS(buffer_size_t) until_turn_ebe, char const* config_str )
{
    std::string const c = node_config( node_id_str, config_str );
    return NS(TrackJob_new_detailed)( "hip", pbuf, num_psets, pset_indices, ebuf, obuf, until_turn_ebe,
                                      c.c_str() );
}

bool NS(CudaTrackJob_has_controller)( const NS(CudaTrackJob)* ) { return false; }
void* NS(CudaTrackJob_get_ptr_controller)( NS(CudaTrackJob)* ) { return nullptr; }
void const* NS(CudaTrackJob_get_ptr_const_controller)( const NS(CudaTrackJob)* ) { return nullptr; }

#define ST_CUDA_TRACK_JOB_ARG_GETTERS( what ) \
    bool NS(CudaTrackJob_has_##what##_arg)( const NS(CudaTrackJob)* ) { return false; } \
    void* NS(CudaTrackJob_get_ptr_##what##_arg)( NS(CudaTrackJob)* ) { return nullptr; } \
    void const* NS(CudaTrackJob_get_ptr_const_##what##_arg)( const NS(CudaTrackJob)* ) { return nullptr; }
ST_CUDA_TRACK_JOB_ARG_GETTERS( particles )
ST_CUDA_TRACK_JOB_ARG_GETTERS( beam_elements )
ST_CUDA_TRACK_JOB_ARG_GETTERS( output )
ST_CUDA_TRACK_JOB_ARG_GETTERS( elem_by_elem_config )
ST_CUDA_TRACK_JOB_ARG_GETTERS( debug_register )
ST_CUDA_TRACK_JOB_ARG_GETTERS( particles_addr )
#undef ST_CUDA_TRACK_JOB_ARG_GETTERS

NS(arch_size_t) NS(CudaTrackJob_default_threads_per_block)( const NS(CudaTrackJob)* j )
{
    return j ? ( NS(arch_size_t) )NS(TrackJob_block_size)( j ) : 0u;
}

NS(arch_size_t) NS(CudaTrackJob_default_track_threads_per_block)( const NS(CudaTrackJob)* j )
{
    return NS(CudaTrackJob_default_threads_per_block)( j );
}

}  // extern "C"

/* sixtracklib_amd/track_job.h -- the track-job C API (drop-in boundary).
 *
 * Replaces the reference's legacy track-job API used by the Python
 * ``sixtracklib.TrackJob`` (sixtracklib/common/track_job.h:139-470,
 * implemented in common/internal/track_job.cpp; ctypes prototypes in
 * python/sixtracklib/stcommon.py:1334-1375).  Names, argument meaning,
 * return conventions and ownership rules are the reference's:
 *   - constructors return NULL on failure;
 *   - track functions return NS(track_status_t): 0 success, -1 failure,
 *     per-particle statuses OR'ed; a NULL job gives -1; particle loss is
 *     not an error;
 *   - the caller owns the particle and beam-element buffers, the job keeps
 *     pointers to them; a NULL output buffer makes the job allocate and own
 *     one when output is needed (element-by-element or BeamMonitors);
 *   - collect writes device results back into the caller's buffers in
 *     place; push uploads host changes.
 * The architecture string is "hip" (also accepted: "amd", "amdgpu",
 * "gpu").  There is no CPU architecture in this library: "cpu" and
 * "opencl" are rejected (constructor returns NULL).
 * config_str tokens (space or comma separated): "fma=0|1" (default 0:
 * results bit-identical to the reference CPU tracker; 1: fused
 * multiply-add build: the energy fields and the path length rounded as
 * the reference rounds them, the transverse coordinates within about
 * 1e-13 of their rms after one turn of a 20,000-element ring (more where
 * the orbit swings far beyond the beam: 1.5e-11 on the reference's LHC
 * lattice, whose 12 mm bumps dwarf a 0.2 mm beam), loss elements and
 * turns identical on regular orbits, those whose exact-build loss does not
 * move when the start moves by one ulp), "block=64..256",
 * "device=N", "ebe_rolling=0|1" (default 1, as the reference job),
 * "compact=0|1" (default 1: track_until first gathers the particles that
 * still have turns to go, so lost ones take no lanes; results identical),
 * "turns_per_launch=K" (default 0: chosen from the work size; with
 * compact=1 a request is split into launches of K turns and the particle
 * list is rebuilt between them; results identical), "queues=1|8" (default
 * 1: one work queue for the persistent launch; 8: one per XCD),
 * "ppl=1|2|auto" (default auto: particles per lane of the until-turn kernel
 * of lean lattices; auto takes two in the FMA build for a request whose
 * particles fill at least 70 % of the resident waves with 128-particle
 * tiles and one otherwise, and between one and two particles per resident
 * lane it runs mixed launches, some waves of every SIMD with two particles
 * per lane and the others with one (the most pair waves that leave one
 * single wave per SIMD); with 2 to 8 waves per SIMD of one particle per
 * lane, the particles spread evenly over that many waves of every SIMD (6
 * of them, 2 pairs, for 8); the mixed launches in both builds; results
 * identical), "mix=0|1" (default 1: the mixed launches of ppl=auto),
 * "balance=0|1" (default 1: with compact=1
 * every launch gives each resident wave one 64-particle tile, the
 * particles furthest behind first, and the host reads the counts back
 * between launches; 0: every chunk hands out all its tiles; results
 * identical), "wave_prio=0|1" (default 1: the waves of a launch are given
 * issue priority by their progress, so that they finish together; results
 * identical), "waves_per_cu=N" (default: the kernel's full occupancy;
 * fewer resident waves of the persistent launch, for tests),
 * "ebe_store_passes=K" (default 0: the element-by-element store holds
 * particles x elements x turns records, as the reference's; K > 0 sizes it
 * for K element passes of every particle when that is fewer, and with
 * ebe_rolling=1 it then keeps each particle's K most recent passes: record
 * (turn t, element e, particle p) goes to slot ((t - t0) * E + e) mod K * P
 * + p, the reference's rolling index with capacity K * P; with
 * ebe_rolling=0 it keeps the first K), "staging=scalar|lds" (default
 * scalar: the until-turn kernel of lean lattices reads the lattice
 * through scalar loads in every wave; lds: 256-lane workgroups copy it
 * piece by piece into LDS and walk it together; one particle per lane;
 * results identical).
 */
#ifndef SIXTRACKLIB_AMD_TRACK_JOB_H__
#define SIXTRACKLIB_AMD_TRACK_JOB_H__

#include "sixtracklib_amd/definitions.h"
#include "sixtracklib_amd/buffer.h"
#include "sixtracklib_amd/particles.h"

#if defined( __cplusplus )
extern "C" {
#endif

typedef struct NS(TrackJobBase) NS(TrackJobBase);
typedef NS(TrackJobBase) NS(TrackJob);

/* Buffer ids of a job (control/definitions.h:134-168). */
#define ST_ARCH_ILLEGAL_BUFFER_ID          ( ( NS(arch_size_t) )0u )
#define ST_ARCH_PARTICLES_BUFFER_ID        ( ( NS(arch_size_t) )1u )
#define ST_ARCH_BEAM_ELEMENTS_BUFFER_ID    ( ( NS(arch_size_t) )2u )
#define ST_ARCH_OUTPUT_BUFFER_ID           ( ( NS(arch_size_t) )3u )
#define ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ( ( NS(arch_size_t) )1024u )

/* buffer/assign_address_item.h:34-45: "write the address of field
 * src_pointer_offset of object src_elem_index (type src_elem_type_id) of
 * buffer src_buffer_id into field dest_pointer_offset of object
 * dest_elem_index (type dest_elem_type_id) of buffer dest_buffer_id". */
typedef struct NS(AssignAddressItem)
{
    NS(buffer_size_t)    dest_buffer_id;
    NS(buffer_size_t)    src_buffer_id;
    NS(object_type_id_t) dest_elem_type_id;
    NS(object_type_id_t) src_elem_type_id;
    NS(buffer_size_t)    dest_elem_index;
    NS(buffer_size_t)    src_elem_index;
    NS(buffer_size_t)    dest_pointer_offset;
    NS(buffer_size_t)    src_pointer_offset;
}
NS(AssignAddressItem);

/* buffer/assign_address_item.h:27-32 */
typedef struct NS(AssignAddressItemDestSrcPair)
{
    NS(buffer_size_t) dest_buffer_id;
    NS(buffer_size_t) src_buffer_id;
}
NS(AssignAddressItemDestSrcPair);
typedef NS(AssignAddressItemDestSrcPair) NS(TrackJobDestSrcBufferIds);

/* output/elem_by_elem_config.h:33-53: the element-by-element store
 * (order: only turn-element-particle, value 0, exists) */
typedef int64_t NS(elem_by_elem_order_t);
#define ST_ELEM_BY_ELEM_ORDER_TURN_ELEM_PARTICLES ( ( NS(elem_by_elem_order_t) )0 )
#define ST_ELEM_BY_ELEM_ORDER_DEFAULT ST_ELEM_BY_ELEM_ORDER_TURN_ELEM_PARTICLES
typedef struct NS(ElemByElemConfig)
{
    NS(elem_by_elem_order_t) order;
    int64_t num_particles_to_store, num_elements_to_store, num_turns_to_store;
    int64_t min_particle_id, min_element_id, min_turn;
    int64_t max_particle_id, max_element_id, max_turn;
    int64_t is_rolling;
    NS(buffer_addr_t) out_store_addr;
}
NS(ElemByElemConfig);

/* track_job.h:123-135: the device id in a configuration string
 * ("device=N", or a leading "P.D" platform.device pair as the reference's
 * OpenCL ids); the lower-case arch string.  0 on success. */
ST_API int NS(TrackJob_extract_device_id_str)( char const* config_str, char* device_id_str,
    NS(buffer_size_t) max_device_id_str_len );
ST_API int NS(TrackJob_sanitize_arch_str_inplace)( char* arch_str, NS(buffer_size_t) max_arch_str_len );
ST_API int NS(TrackJob_sanitize_arch_str)( char const* arch_str, char* sanitized_arch_str,
    NS(buffer_size_t) max_arch_str_len );

/* track_job.h:139 */
ST_API NS(TrackJobBase)* NS(TrackJob_create)( char const* arch, char const* config_str );
/* track_job.h:143 */
ST_API NS(TrackJobBase)* NS(TrackJob_new)( char const* arch,
    NS(Buffer)* particles_buffer, NS(Buffer)* beam_elements_buffer,
    char const* config_str );
/* track_job.h:149 */
ST_API NS(TrackJobBase)* NS(TrackJob_new_with_output)( char const* arch,
    NS(Buffer)* particles_buffer, NS(Buffer)* beam_elements_buffer,
    NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem,
    char const* config_str );
/* track_job.h:157 */
ST_API NS(TrackJobBase)* NS(TrackJob_new_detailed)( char const* arch,
    NS(Buffer)* particles_buffer, NS(buffer_size_t) num_particle_sets,
    NS(buffer_size_t) const* particle_set_indices_begin,
    NS(Buffer)* beam_elements_buffer, NS(Buffer)* output_buffer,
    NS(buffer_size_t) until_turn_elem_by_elem, char const* config_str );
/* track_job.h:169 */
ST_API void NS(TrackJob_delete)( NS(TrackJobBase)* job );

/* track_job.h:172-184 */
ST_API NS(track_status_t) NS(TrackJob_track_until)( NS(TrackJobBase)* job,
    NS(buffer_size_t) until_turn );
ST_API NS(track_status_t) NS(TrackJob_track_elem_by_elem)( NS(TrackJobBase)* job,
    NS(buffer_size_t) until_turn );
ST_API NS(track_status_t) NS(TrackJob_track_line)( NS(TrackJobBase)* job,
    NS(buffer_size_t) begin_idx, NS(buffer_size_t) end_idx, bool finish_turn );

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory while they track (tight serving loops, code that must not
 * call hipMalloc): NS(TrackJob_track_until) with the same arguments,
 * scheduling, kernels and results, but every scratch array of the call
 * (status word, particle-set descriptors, counts, work queues,
 * active-particle lists) lives in one static device array of the library
 * (8.06 MiB on each GPU the library runs on, held until the process
 * ends); the call allocates and frees nothing.  Limits: at most 64
 * particle sets of together at most 2,097,152 particles (status -1 and a
 * message in NS(Hip_last_error_str) beyond them).  The job itself (its
 * particle, lattice and output copies) is made by the usual constructors.
 * All calls on one GPU share that array, which keeps the last call's
 * contents: calls that overlap in time on one GPU (from several host
 * threads) race and must not be made. */
ST_API NS(track_status_t) NS(TrackJob_track_until_noalloc)( NS(TrackJobBase)* job,
    NS(buffer_size_t) until_turn );

/* track_job.h:188-244 */
ST_API void NS(TrackJob_collect)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_collect_detailed)( NS(TrackJobBase)* job, NS(track_job_collect_flag_t) flags );
ST_API void NS(TrackJob_collect_particles)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_collect_beam_elements)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_collect_output)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_enable_collect_particles)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_disable_collect_particles)( NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_is_collecting_particles)( const NS(TrackJobBase)* job );
ST_API void NS(TrackJob_enable_collect_beam_elements)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_disable_collect_beam_elements)( NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_is_collecting_beam_elements)( const NS(TrackJobBase)* job );
ST_API void NS(TrackJob_enable_collect_output)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_disable_collect_output)( NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_is_collecting_output)( const NS(TrackJobBase)* job );
ST_API NS(track_job_collect_flag_t) NS(TrackJob_get_collect_flags)( const NS(TrackJobBase)* job );
ST_API void NS(TrackJob_set_collect_flags)( NS(TrackJobBase)* job, NS(track_job_collect_flag_t) flags );
ST_API bool NS(TrackJob_requires_collecting)( const NS(TrackJobBase)* job );

/* track_job.h:248-262 */
ST_API void NS(TrackJob_push)( NS(TrackJobBase)* job, NS(track_job_push_flag_t) flags );
ST_API void NS(TrackJob_push_particles)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_push_beam_elements)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_push_output)( NS(TrackJobBase)* job );

/* track_job.h:206-290: particle addresses.  A job keeps one
 * NS(ParticlesAddr) per Particles object of its particle buffer; fetch
 * fills them with the device addresses of the job's copy (valid until the
 * job is reset or deleted), clear zeroes them.  They are made on the host:
 * collect and push of the addresses have nothing to move. */
ST_API void NS(TrackJob_collect_particles_addresses)( NS(TrackJobBase)* job );
ST_API void NS(TrackJob_push_particles_addresses)( NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_can_fetch_particles_addr)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_has_particles_addr)( const NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_fetch_particle_addresses)( NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_fetch_particles_addr)( NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_clear_particle_addresses)( NS(TrackJobBase)* job,
    NS(arch_size_t) index );
ST_API NS(arch_status_t) NS(TrackJob_clear_particles_addr)( NS(TrackJobBase)* job,
    NS(buffer_size_t) particle_set_index );
ST_API NS(arch_status_t) NS(TrackJob_clear_all_particle_addresses)( NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_clear_all_particles_addr)( NS(TrackJobBase)* job );
ST_API NS(ParticlesAddr) const* NS(TrackJob_particle_addresses)( const NS(TrackJobBase)* job,
    NS(arch_size_t) index );
ST_API NS(ParticlesAddr) const* NS(TrackJob_particles_addr)( const NS(TrackJobBase)* job,
    NS(buffer_size_t) particle_set_index );
ST_API NS(Buffer)* NS(TrackJob_get_particles_addr_buffer)( NS(TrackJobBase)* job );
ST_API NS(Buffer) const* NS(TrackJob_get_const_particles_addr_buffer)( const NS(TrackJobBase)* job );
ST_API NS(Buffer) const* NS(TrackJob_particles_addr_buffer)( const NS(TrackJobBase)* job );

/* track_job.h:300-336 */
ST_API void NS(TrackJob_clear)( NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_reset)( NS(TrackJobBase)* job, NS(Buffer)* particles_buffer,
    NS(Buffer)* beam_elements_buffer, NS(Buffer)* output_buffer );
ST_API bool NS(TrackJob_reset_with_output)( NS(TrackJobBase)* job,
    NS(Buffer)* particles_buffer, NS(Buffer)* beam_elements_buffer,
    NS(Buffer)* output_buffer, NS(buffer_size_t) until_turn_elem_by_elem );
ST_API bool NS(TrackJob_reset_detailed)( NS(TrackJobBase)* job,
    NS(Buffer)* particles_buffer, NS(buffer_size_t) num_particle_sets,
    NS(buffer_size_t) const* particle_set_indices_begin,
    NS(Buffer)* beam_elements_buffer, NS(Buffer)* output_buffer,
    NS(buffer_size_t) until_turn_elem_by_elem );
ST_API bool NS(TrackJob_reset_particle_set)( NS(TrackJobBase)* job,
    NS(Buffer)* particles_buffer, NS(buffer_size_t) particle_set_index,
    NS(Buffer)* beam_elements_buffer, NS(Buffer)* output_buffer );
ST_API bool NS(TrackJob_select_particle_set)( NS(TrackJobBase)* job,
    NS(buffer_size_t) particle_set_index );
ST_API bool NS(TrackJob_assign_output_buffer)( NS(TrackJobBase)* job,
    NS(Buffer)* output_buffer );

/* track_job.h:342-470 introspection */
ST_API NS(track_job_type_t) NS(TrackJob_get_type_id)( const NS(TrackJobBase)* job );
ST_API char const* NS(TrackJob_get_type_str)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_has_device_id_str)( const NS(TrackJobBase)* job );
ST_API char const* NS(TrackJob_get_device_id_str)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_has_config_str)( const NS(TrackJobBase)* job );
ST_API char const* NS(TrackJob_get_config_str)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_num_particle_sets)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) const* NS(TrackJob_get_particle_set_indices_begin)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) const* NS(TrackJob_get_particle_set_indices_end)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_particle_set_index)( const NS(TrackJobBase)* job, NS(buffer_size_t) n );
ST_API NS(buffer_size_t) NS(TrackJob_get_total_num_particles_in_sets)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) const* NS(TrackJob_get_num_particles_in_sets_begin)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) const* NS(TrackJob_get_num_particles_in_sets_end)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_num_particles_in_set)( const NS(TrackJobBase)* job,
    NS(buffer_size_t) n );
ST_API NS(particle_index_t) NS(TrackJob_get_min_particle_id)( const NS(TrackJobBase)* job );
ST_API NS(particle_index_t) NS(TrackJob_get_max_particle_id)( const NS(TrackJobBase)* job );
ST_API NS(particle_index_t) NS(TrackJob_get_min_element_id)( const NS(TrackJobBase)* job );
ST_API NS(particle_index_t) NS(TrackJob_get_max_element_id)( const NS(TrackJobBase)* job );
ST_API NS(particle_index_t) NS(TrackJob_get_min_initial_turn_id)( const NS(TrackJobBase)* job );
ST_API NS(particle_index_t) NS(TrackJob_get_max_initial_turn_id)( const NS(TrackJobBase)* job );
ST_API NS(Buffer)* NS(TrackJob_get_particles_buffer)( NS(TrackJobBase)* job );
ST_API NS(Buffer) const* NS(TrackJob_get_const_particles_buffer)( const NS(TrackJobBase)* job );
ST_API NS(Buffer)* NS(TrackJob_get_beam_elements_buffer)( NS(TrackJobBase)* job );
ST_API NS(Buffer) const* NS(TrackJob_get_const_beam_elements_buffer)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_has_output_buffer)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_owns_output_buffer)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_has_elem_by_elem_output)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_has_beam_monitor_output)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_beam_monitor_output_buffer_offset)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_elem_by_elem_output_buffer_offset)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_num_elem_by_elem_turns)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_num_beam_monitors)( const NS(TrackJobBase)* job );
ST_API NS(Buffer)* NS(TrackJob_get_output_buffer)( NS(TrackJobBase)* job );
ST_API NS(Buffer) const* NS(TrackJob_get_const_output_buffer)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_has_beam_monitors)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) const* NS(TrackJob_get_beam_monitor_indices_begin)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) const* NS(TrackJob_get_beam_monitor_indices_end)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_get_beam_monitor_index)( const NS(TrackJobBase)* job,
    NS(buffer_size_t) n );
/* the element-by-element configuration of the job's output (NULL when
 * the job stores no element-by-element output); out_store_addr is the
 * host address of the store in the output buffer */
ST_API bool NS(TrackJob_has_elem_by_elem_config)( const NS(TrackJobBase)* job );
ST_API NS(ElemByElemConfig) const* NS(TrackJob_get_elem_by_elem_config)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_is_elem_by_elem_config_rolling)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_get_default_elem_by_elem_config_rolling_flag)( const NS(TrackJobBase)* job );
/* takes effect at the next reset, as the reference's */
ST_API void NS(TrackJob_set_default_elem_by_elem_config_rolling_flag)( NS(TrackJobBase)* job,
    bool is_rolling_flag );
ST_API NS(elem_by_elem_order_t) NS(TrackJob_get_elem_by_elem_config_order)( const NS(TrackJobBase)* job );
ST_API NS(elem_by_elem_order_t) NS(TrackJob_get_default_elem_by_elem_config_order)(
    const NS(TrackJobBase)* job );
ST_API void NS(TrackJob_set_default_elem_by_elem_config_order)( NS(TrackJobBase)* job,
    NS(elem_by_elem_order_t) order );
ST_API NS(track_status_t) NS(TrackJob_last_status)( const NS(TrackJobBase)* job );

/* Timing and build-variant introspection (no reference counterpart):
 * device time of the kernels of the last track_* call, measured with HIP
 * events on the job's stream; whether the FMA build and the full-element
 * kernel (beam-beam, space charge, monitors, RF multipoles) are used. */
ST_API double NS(TrackJob_last_kernel_time_ms)( const NS(TrackJobBase)* job );
ST_API int NS(TrackJob_uses_fma)( const NS(TrackJobBase)* job );
ST_API int NS(TrackJob_uses_full_kernel)( const NS(TrackJobBase)* job );
/* threads per block of the job's launches (config token "block=") */
ST_API int NS(TrackJob_block_size)( const NS(TrackJobBase)* job );

/* Output-buffer sizing used by the Python binding before it builds its
 * own output buffer (output/output_buffer.c:44-131 + 288-363):
 * reports the particle-set sizes the job will need, in the order the
 * output buffer must hold them (element-by-element set first, then one
 * per BeamMonitor).  config_str (may be NULL) is the job's configuration,
 * of which "ebe_store_passes" changes the sizes.  Returns the number of
 * sets, -1 on error. */
ST_API int64_t NS(TrackJob_required_output_sets)( NS(Buffer)* particles_buffer,
    NS(buffer_size_t) num_particle_sets,
    NS(buffer_size_t) const* particle_set_indices_begin,
    NS(Buffer)* beam_elements_buffer, NS(buffer_size_t) until_turn_elem_by_elem,
    int64_t* out_sizes, int64_t max_sets, int64_t* has_elem_by_elem,
    char const* config_str );

/* Stored buffers (track_job.h:717-772).  A stored buffer is a relocated
 * CObjects buffer the job keeps beside its particles, beam elements and
 * output (ids from ST_ARCH_MIN_USER_DEFINED_BUFFER_ID on, stable when a
 * buffer is removed).  On this device its contents reach the kernels as
 * copies inside the packed lattice: the line-density tables of
 * SCInterpolatedProfile elements whose interpol_data_addr points into it.
 * push rebuilds those copies; collect has nothing to bring back.
 * take_ownership: the job deletes the buffer (NS(Buffer_delete)) when it
 * is removed or the job is deleted; delete_ptr_after_move is accepted and
 * ignored (no buffer is moved). */
ST_API NS(arch_size_t) NS(TrackJob_add_stored_buffer)( NS(TrackJobBase)* job,
    NS(Buffer)* buffer, bool take_ownership, bool delete_ptr_after_move );
ST_API NS(arch_size_t) NS(TrackJob_create_stored_buffer)( NS(TrackJobBase)* job,
    NS(buffer_size_t) buffer_capacity );
ST_API NS(arch_status_t) NS(TrackJob_remove_stored_buffer)( NS(TrackJobBase)* job,
    NS(arch_size_t) buffer_id );
ST_API bool NS(TrackJob_has_stored_buffers)( const NS(TrackJobBase)* job );
ST_API NS(arch_size_t) NS(TrackJob_num_stored_buffers)( const NS(TrackJobBase)* job );
ST_API NS(arch_size_t) NS(TrackJob_min_stored_buffer_id)( const NS(TrackJobBase)* job );
ST_API NS(arch_size_t) NS(TrackJob_max_stored_buffer_id)( const NS(TrackJobBase)* job );
ST_API bool NS(TrackJob_owns_stored_buffer)( const NS(TrackJobBase)* job, NS(arch_size_t) buffer_id );
ST_API NS(Buffer)* NS(TrackJob_stored_buffer)( NS(TrackJobBase)* job, NS(arch_size_t) buffer_id );
ST_API NS(Buffer) const* NS(TrackJob_const_stored_buffer)( const NS(TrackJobBase)* job,
    NS(arch_size_t) buffer_id );
ST_API NS(arch_size_t) NS(TrackJob_stored_buffers_capacity)( const NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_reserve_stored_buffers_capacity)( NS(TrackJobBase)* job,
    NS(arch_size_t) capacity );
/* buffer ids 1 (particles), 2 (beam elements), 3 (output), 1024 + k
 * (stored buffers); no id names raw memory in this library */
ST_API NS(Buffer)* NS(TrackJob_buffer_by_buffer_id)( NS(TrackJobBase)* job, NS(arch_size_t) buffer_id );
ST_API NS(Buffer) const* NS(TrackJob_const_buffer_by_buffer_id)( const NS(TrackJobBase)* job,
    NS(arch_size_t) buffer_id );
ST_API bool NS(TrackJob_is_buffer_by_buffer_id)( const NS(TrackJobBase)* job, NS(arch_size_t) buffer_id );
ST_API bool NS(TrackJob_is_raw_memory_by_buffer_id)( const NS(TrackJobBase)* job,
    NS(arch_size_t) buffer_id );
ST_API NS(arch_status_t) NS(TrackJob_push_stored_buffer)( NS(TrackJobBase)* job,
    NS(arch_size_t) buffer_id );
ST_API NS(arch_status_t) NS(TrackJob_collect_stored_buffer)( NS(TrackJobBase)* job,
    NS(arch_size_t) buffer_id );

/* Address assignment items (track_job.h:500-660).  Items are kept in the
 * order added (an identical item is not added twice).  assign_* writes the
 * host addresses into the host buffers and, when beam elements changed,
 * rebuilds the device lattice.  Destinations other than the beam-element
 * buffer are rejected (status -1, message in NS(Hip_last_error_str)): the
 * job prepares its own output addresses. */
ST_API NS(AssignAddressItem)* NS(TrackJob_add_assign_address_item)( NS(TrackJobBase)* job,
    const NS(AssignAddressItem)* assign_item_to_add );
ST_API NS(AssignAddressItem)* NS(TrackJob_add_assign_address_item_detailed)(
    NS(TrackJobBase)* job, NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id,
    NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset );
/* the reference header also declares this name (track_job.h:631) */
ST_API NS(AssignAddressItem)* NS(TrackJob_add_assign_address_item_detail)( NS(TrackJobBase)* job,
    NS(object_type_id_t) dest_type_id, NS(arch_size_t) dest_buffer_id,
    NS(arch_size_t) dest_elem_index, NS(arch_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(arch_size_t) src_buffer_id,
    NS(arch_size_t) src_elem_index, NS(arch_size_t) src_pointer_offset );
ST_API NS(arch_status_t) NS(TrackJob_remove_assign_address_item)( NS(TrackJobBase)* job,
    NS(AssignAddressItem)* item_to_remove );
ST_API NS(arch_status_t) NS(TrackJob_remove_assign_address_item_by_key_and_index)(
    NS(TrackJobBase)* job, const NS(TrackJobDestSrcBufferIds)* key, NS(buffer_size_t) index );
ST_API bool NS(TrackJob_has_assign_address_item)( const NS(TrackJobBase)* job,
    const NS(AssignAddressItem)* item );
ST_API bool NS(TrackJob_has_assign_item_by_index)( const NS(TrackJobBase)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) index );
ST_API bool NS(TrackJob_has_assign_address_item_detailed)( const NS(TrackJobBase)* job,
    NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id,
    NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset );
/* index among the items of its (dest, src) pair; the pair's item count
 * when there is no such item */
ST_API NS(buffer_size_t) NS(TrackJob_index_of_assign_address_item_detailed)(
    const NS(TrackJobBase)* job,
    NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id,
    NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset );
ST_API NS(buffer_size_t) NS(TrackJob_index_of_assign_address_item)( const NS(TrackJobBase)* job,
    const NS(AssignAddressItem)* item );
ST_API NS(AssignAddressItem) const* NS(TrackJob_ptr_assign_address_item)( const NS(TrackJobBase)* job,
    const NS(AssignAddressItem)* item );
ST_API NS(AssignAddressItem) const* NS(TrackJob_ptr_assign_address_item_detailed)(
    const NS(TrackJobBase)* job,
    NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id,
    NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset );
ST_API NS(AssignAddressItem) const* NS(TrackJob_ptr_assign_address_item_by_index)(
    const NS(TrackJobBase)* job, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) src_buffer_id, NS(buffer_size_t) index );
/* distinct (dest, src) pairs among the items, in the order first added */
ST_API NS(buffer_size_t) NS(TrackJob_num_distinct_available_assign_address_items_dest_src_pairs)(
    const NS(TrackJobBase)* job );
ST_API NS(arch_size_t) NS(TrackJob_num_distinct_assign_items_dest_src_pairs)( const NS(TrackJobBase)* job );
ST_API NS(buffer_size_t) NS(TrackJob_available_assign_address_items_dest_src_pairs)(
    const NS(TrackJobBase)* job, NS(buffer_size_t) max_num_pairs,
    NS(TrackJobDestSrcBufferIds)* pairs_begin );
ST_API NS(arch_size_t) NS(TrackJob_assign_items_dest_src_pairs)( const NS(TrackJobBase)* job,
    NS(arch_size_t) max_num_pairs, NS(AssignAddressItemDestSrcPair)* pairs_begin );
ST_API NS(TrackJobDestSrcBufferIds) const* NS(TrackJob_assign_item_dest_src_begin)(
    const NS(TrackJobBase)* job );
ST_API NS(TrackJobDestSrcBufferIds) const* NS(TrackJob_assign_item_dest_src_end)(
    const NS(TrackJobBase)* job );
/* the items of one (dest, src) pair as CObjects (type 65536, one
 * NS(AssignAddressItem) each) of a buffer the job rebuilds at each call;
 * valid until the next call or change of the items */
ST_API NS(Object) const* NS(TrackJob_assign_items_begin)( const NS(TrackJobBase)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );
ST_API NS(Object) const* NS(TrackJob_assign_items_end)( const NS(TrackJobBase)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );
/* the items of one (dest, src) pair lie next to each other: [begin, end) */
ST_API NS(AssignAddressItem) const* NS(TrackJob_assign_items_dest_src_begin)(
    const NS(TrackJobBase)* job, NS(arch_size_t) dest_buffer_id, NS(arch_size_t) src_buffer_id );
ST_API NS(AssignAddressItem) const* NS(TrackJob_assign_items_dest_src_end)(
    const NS(TrackJobBase)* job, NS(arch_size_t) dest_buffer_id, NS(arch_size_t) src_buffer_id );
ST_API bool NS(TrackJob_has_assign_items)( const NS(TrackJobBase)* job,
    NS(arch_size_t) dest_buffer_id, NS(arch_size_t) src_buffer_id );
ST_API NS(arch_size_t) NS(TrackJob_num_assign_items)( const NS(TrackJobBase)* job,
    NS(arch_size_t) dest_buffer_id, NS(arch_size_t) src_buffer_id );
ST_API NS(arch_size_t) NS(TrackJob_total_num_assign_items)( const NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_commit_address_assignments)( NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_assign_all_addresses)( NS(TrackJobBase)* job );
ST_API NS(arch_status_t) NS(TrackJob_assign_addresses)( NS(TrackJobBase)* job,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id );

/* Device/runtime introspection (no reference counterpart). */
ST_API int NS(Hip_num_devices)( void );
ST_API char const* NS(Hip_last_error_str)( void );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_TRACK_JOB_H__ */

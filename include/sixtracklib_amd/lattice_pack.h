/* sixtracklib_amd/lattice_pack.h -- flatten a beam-element buffer into the
 * packed device stream the kernels read.
 *
 * No counterpart in the reference (whose kernels walk the CObjects buffer
 * directly, track/track.h:209-585); exported for the batched and adjoint
 * front ends and for tests.
 */
#ifndef SIXTRACKLIB_AMD_LATTICE_PACK_H__
#define SIXTRACKLIB_AMD_LATTICE_PACK_H__

#include "sixtracklib_amd/definitions.h"

#if defined( __cplusplus )
extern "C" {
#endif

/* ``belems``: a beam-element CObjects image relocated to its own address.
 * Returns the number of 8-byte words of the packed stream, or -1 on error
 * (unrelocated buffer, element data outside the buffer, ...).  With
 * ``out`` == NULL only sizes are reported.  The stream ends with an end
 * record (2 words).  ``rec_offsets`` (optional, n_elem + 1 entries)
 * receives the word offset of every element record and, last, of the end
 * record;
 * ``monitor_fields`` (optional, one per BeamMonitor) the word index of the
 * monitor's output-address slot. */
ST_API int64_t NS(Lattice_pack)( const unsigned char* belems, uint64_t* out,
    int64_t out_capacity, uint32_t* rec_offsets, int64_t* n_elem_out,
    uint32_t* monitor_fields, int64_t* n_monitors_out );

/* As NS(Lattice_pack), with ``n_table_buffers`` further relocated CObjects
 * images (``table_buffers``) in which the line-density table of an
 * SCInterpolatedProfile (interpol_data_addr) may lie: the table is copied
 * into the element's record when the profile object and its two arrays
 * lie inside one buffer, the lattice's own or one of these. */
ST_API int64_t NS(Lattice_pack_with_tables)( const unsigned char* belems,
    const unsigned char* const* table_buffers, int64_t n_table_buffers, uint64_t* out,
    int64_t out_capacity, uint32_t* rec_offsets, int64_t* n_elem_out,
    uint32_t* monitor_fields, int64_t* n_monitors_out );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_LATTICE_PACK_H__ */

/* sixtracklib_amd/buffer.h -- CObjects managed buffers (host side).
 *
 * Drop-in for the reference's NS(Buffer) C API as used by the track-job
 * path and by the Python bindings (python/sixtracklib/stcommon.py).  Each
 * declaration names the reference declaration it replaces.  The binary
 * layout is the reference format (buffer/managed_buffer_minimal.h): files
 * written by either library load in the other.
 */
#ifndef SIXTRACKLIB_AMD_BUFFER_H__
#define SIXTRACKLIB_AMD_BUFFER_H__

#include "sixtracklib_amd/definitions.h"

#if defined( __cplusplus )
extern "C" {
#endif

/* buffer/buffer_object.h:24-30 */
typedef struct NS(Object)
{
    NS(buffer_addr_t)     begin_addr;
    NS(object_type_id_t)  type_id;
    NS(buffer_size_t)     size;
}
NS(Object);

/* Opaque; the reference struct is buffer/buffer_type.h (NS(Buffer)). */
typedef struct NS(Buffer) NS(Buffer);

/* common/buffer.h: NS(Buffer_new) */
ST_API NS(Buffer)* NS(Buffer_new)( NS(buffer_size_t) buffer_capacity );
/* common/buffer.h: NS(Buffer_new_detailed) */
ST_API NS(Buffer)* NS(Buffer_new_detailed)( NS(buffer_size_t) max_num_objects,
    NS(buffer_size_t) max_num_slots, NS(buffer_size_t) max_num_dataptrs,
    NS(buffer_size_t) max_num_garbage_ranges, uint64_t buffer_flags );
/* common/buffer.h: NS(Buffer_new_from_file) */
ST_API NS(Buffer)* NS(Buffer_new_from_file)( char const* path );
/* common/buffer.h: NS(Buffer_new_on_memory) / Python
 * st_Buffer_new_mapped_on_cbuffer (stcommon.py): views caller-owned memory
 * (8-byte aligned) without copying; relocates it in place if needed. */
ST_API NS(Buffer)* NS(Buffer_new_on_memory)( unsigned char* data,
    NS(buffer_size_t) size );
ST_API NS(Buffer)* NS(Buffer_new_mapped_on_cbuffer)( unsigned char* data,
    NS(buffer_size_t) size );
/* common/buffer.h: NS(Buffer_new_from_copy) */
ST_API NS(Buffer)* NS(Buffer_new_from_copy)( const NS(Buffer)* other );
/* common/buffer.h: NS(Buffer_delete) */
ST_API void NS(Buffer_delete)( NS(Buffer)* buffer );

/* common/buffer.h: NS(Buffer_write_to_file) */
ST_API bool NS(Buffer_write_to_file)( const NS(Buffer)* buffer, char const* path );
/* common/buffer.h: NS(Buffer_needs_remapping) / NS(Buffer_remap) */
ST_API bool NS(Buffer_needs_remapping)( const NS(Buffer)* buffer );
ST_API int  NS(Buffer_remap)( NS(Buffer)* buffer );
/* common/buffer.h: NS(Buffer_clear) (keeps capacity) */
ST_API int  NS(Buffer_clear)( NS(Buffer)* buffer, bool set_data_to_zero );
/* common/buffer.h: NS(Buffer_reserve) */
ST_API int  NS(Buffer_reserve)( NS(Buffer)* buffer,
    NS(buffer_size_t) max_num_objects, NS(buffer_size_t) max_num_slots,
    NS(buffer_size_t) max_num_dataptrs, NS(buffer_size_t) max_num_garbage );

ST_API NS(buffer_size_t) NS(Buffer_get_num_of_objects)( const NS(Buffer)* b );
ST_API NS(buffer_size_t) NS(Buffer_get_size)( const NS(Buffer)* b );
ST_API NS(buffer_size_t) NS(Buffer_get_capacity)( const NS(Buffer)* b );
ST_API NS(buffer_size_t) NS(Buffer_get_slot_size)( const NS(Buffer)* b );
ST_API NS(buffer_size_t) NS(Buffer_get_num_of_slots)( const NS(Buffer)* b );
ST_API NS(buffer_size_t) NS(Buffer_get_num_of_dataptrs)( const NS(Buffer)* b );
ST_API NS(buffer_addr_t) NS(Buffer_get_data_begin_addr)( const NS(Buffer)* b );
ST_API unsigned char*    NS(Buffer_get_data_begin)( NS(Buffer)* b );
ST_API unsigned char const* NS(Buffer_get_const_data_begin)( const NS(Buffer)* b );
ST_API bool NS(Buffer_owns_datastore)( const NS(Buffer)* b );

ST_API NS(Object)* NS(Buffer_get_objects_begin)( NS(Buffer)* b );
ST_API NS(Object)* NS(Buffer_get_objects_end)( NS(Buffer)* b );
ST_API NS(Object)* NS(Buffer_get_object)( NS(Buffer)* b, NS(buffer_size_t) index );
ST_API NS(Object) const* NS(Buffer_get_const_object)( const NS(Buffer)* b,
    NS(buffer_size_t) index );

/* common/buffer.h: NS(Buffer_add_object).  Pointer fields of the object
 * struct at byte offsets ``offsets[i]`` receive the address of a fresh
 * array of ``counts[i]`` elements of ``sizes[i]`` bytes, copied from the
 * pointer stored in ``object_handle`` when it is not NULL, else zeroed. */
ST_API NS(Object)* NS(Buffer_add_object)( NS(Buffer)* b,
    const void* object_handle, NS(buffer_size_t) object_size,
    NS(object_type_id_t) type_id, NS(buffer_size_t) num_dataptrs,
    NS(buffer_size_t) const* offsets, NS(buffer_size_t) const* sizes,
    NS(buffer_size_t) const* counts );

/* buffer/buffer_object.h accessors */
ST_API NS(object_type_id_t) NS(Object_get_type_id)( const NS(Object)* obj );
ST_API NS(buffer_addr_t) NS(Object_get_begin_addr)( const NS(Object)* obj );
ST_API NS(buffer_size_t) NS(Object_get_size)( const NS(Object)* obj );

/* Relocate a raw CObjects image in place to its current address
 * (buffer/managed_buffer_remap.h:644-681, NS(ManagedBuffer_remap));
 * ``offset_base`` lets a caller relocate an image to an address other
 * than where it lies (e.g. a device copy); pass 0 to use the image's own
 * address.  Returns 0 on success. */
ST_API int NS(ManagedBuffer_remap_to)( unsigned char* data,
    NS(buffer_addr_t) target_base );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_BUFFER_H__ */

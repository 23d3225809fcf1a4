/* sixtracklib_amd/particles.h -- the SoA particle set.
 *
 * NS(Particles) is layout-identical to the reference struct
 * (sixtracklib/common/particles.h:38-67): a count followed by 17 FP64 and
 * 4 int64 field pointers into the owning CObjects buffer.
 */
#ifndef SIXTRACKLIB_AMD_PARTICLES_H__
#define SIXTRACKLIB_AMD_PARTICLES_H__

#include "sixtracklib_amd/definitions.h"
#include "sixtracklib_amd/buffer.h"

#if defined( __cplusplus )
extern "C" {
#endif

typedef struct NS(Particles)
{
    NS(particle_num_elements_t) num_particles;
    double *q0, *mass0, *beta0, *gamma0, *p0c, *s, *x, *y, *px, *py, *zeta,
           *psigma, *delta, *rpp, *rvv, *chi, *charge_ratio;
    int64_t *particle_id, *at_element_id, *at_turn, *state;
}
NS(Particles);

#define ST_PARTICLES_NUM_REAL_FIELDS  17
#define ST_PARTICLES_NUM_INDEX_FIELDS 4

/* particles/particles_addr.h:27-55: the addresses of the 21 field arrays
 * of one particle set (object type ST_OBJECT_TYPE_PARTICLES_ADDR).  A track
 * job fills them with the addresses of its device copy. */
typedef struct NS(ParticlesAddr)
{
    NS(particle_num_elements_t) num_particles;
    NS(buffer_addr_t) q0_addr, mass0_addr, beta0_addr, gamma0_addr, p0c_addr;
    NS(buffer_addr_t) s_addr, x_addr, y_addr, px_addr, py_addr, zeta_addr;
    NS(buffer_addr_t) psigma_addr, delta_addr, rpp_addr, rvv_addr, chi_addr, charge_ratio_addr;
    NS(buffer_addr_t) particle_id_addr, at_element_id_addr, at_turn_addr, state_addr;
}
NS(ParticlesAddr);

/* particles.h: NS(Particles_new) -- zero-initialised (state = 0). */
ST_API NS(Particles)* NS(Particles_new)( NS(Buffer)* buffer,
    NS(particle_num_elements_t) num_particles );
/* particles.h: NS(Particles_add_copy) */
ST_API NS(Particles)* NS(Particles_add_copy)( NS(Buffer)* buffer,
    const NS(Particles)* orig );
/* particles.h: NS(Particles_buffer_get_particles) */
ST_API NS(Particles)* NS(Particles_buffer_get_particles)( NS(Buffer)* buffer,
    NS(buffer_size_t) index );
ST_API NS(Particles) const* NS(Particles_buffer_get_const_particles)(
    const NS(Buffer)* buffer, NS(buffer_size_t) index );
/* particles.h: NS(Particles_get_num_of_particles) */
ST_API NS(particle_num_elements_t) NS(Particles_get_num_of_particles)(
    const NS(Particles)* p );
/* particles.h: NS(Particles_buffer_get_num_of_particle_blocks) */
ST_API NS(buffer_size_t) NS(Particles_buffer_get_num_of_particle_blocks)(
    const NS(Buffer)* buffer );
/* particles.h: NS(Particles_buffer_get_total_num_of_particles) */
ST_API NS(particle_num_elements_t)
NS(Particles_buffer_get_total_num_of_particles)( const NS(Buffer)* buffer );
/* particles.h: NS(Particles_copy) (same size required) */
ST_API int NS(Particles_copy)( NS(Particles)* dest, const NS(Particles)* src );
/* particles.h:188 NS(Particles_find_min_max_attributes): folds the
 * particle ids, at_element and at_turn of ``p`` into the six running
 * extrema (initialise them with INT64_MAX / -1 or call with fresh values). */
ST_API int NS(Particles_find_min_max_attributes)( const NS(Particles)* p,
    NS(particle_index_t)* min_pid, NS(particle_index_t)* max_pid,
    NS(particle_index_t)* min_elem, NS(particle_index_t)* max_elem,
    NS(particle_index_t)* min_turn, NS(particle_index_t)* max_turn );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_PARTICLES_H__ */

/* oracle/oracle.h -- CPU restatement of the reference tracker.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product (sixtracklib_amd, libsixtrack_amd.so) never links,
 * loads or calls it.
 *
 * The oracle works directly on CObjects buffers that have been relocated to
 * their current host address (oracle_remap), exactly like the reference's
 * CPU path: particles are NS(Particles) objects (count + 21 field pointers,
 * sixtracklib/common/particles.h:38-67), the lattice is the object index of
 * a beam-elements buffer (sixtracklib/common/buffer/buffer_object.h:24-30).
 *
 * Two builds exist:
 *   liboracle.so      transcendentals from include/sixtracklib_amd/stmath.h
 *                     (bit-identical host/device by construction) -- the
 *                     build the GPU kernels are compared against bit-for-bit;
 *   liboracle_libm.so transcendentals from the platform libm, exactly as the
 *                     reference calls them -- the build pinned against the
 *                     reference's golden fixtures.
 * Both are compiled with -ffp-contract=off (the reference is built with
 * -ansi, i.e. no FP contraction; sixtracklib/Settings.cmake.default).
 */
#ifndef SIXTRACKLIB_AMD_ORACLE_H__
#define SIXTRACKLIB_AMD_ORACLE_H__

#include <stdint.h>

typedef struct or_particles_s
{
    int64_t num_particles;
    double *q0, *mass0, *beta0, *gamma0, *p0c, *s, *x, *y, *px, *py, *zeta,
           *psigma, *delta, *rpp, *rvv, *chi, *charge_ratio;
    int64_t *particle_id, *at_element_id, *at_turn, *state;
}
or_particles_t;

/* layout of NS(ElemByElemConfig), output/elem_by_elem_config.h:39-53 */
typedef struct or_ebe_config_s
{
    int64_t  order;
    int64_t  num_particles_to_store;
    int64_t  num_elements_to_store;
    int64_t  num_turns_to_store;
    int64_t  min_particle_id;
    int64_t  min_element_id;
    int64_t  min_turn;
    int64_t  max_particle_id;
    int64_t  max_element_id;
    int64_t  max_turn;
    int64_t  is_rolling;
    uint64_t out_store_addr;
}
or_ebe_config_t;

#ifdef __cplusplus
extern "C" {
#endif

int      oracle_remap( unsigned char* buffer );
int64_t  oracle_num_objects( const unsigned char* buffer );
uint64_t oracle_object_type( const unsigned char* buffer, int64_t index );
void*    oracle_object_begin( const unsigned char* buffer, int64_t index );

int32_t oracle_track_until( or_particles_t* p, const unsigned char* belems,
                            int64_t until_turn, int num_threads );
int32_t oracle_track_elem_by_elem( or_particles_t* p,
                                   const or_ebe_config_t* config,
                                   const unsigned char* belems,
                                   int64_t until_turn );
int32_t oracle_track_line( or_particles_t* p, const unsigned char* belems,
                           int64_t begin_idx, int64_t end_idx,
                           int finish_turn );
int32_t oracle_track_element( or_particles_t* p, int64_t particle_idx,
                              const unsigned char* belems,
                              int64_t element_idx );
int64_t oracle_ebe_store_index( const or_ebe_config_t* config,
                                int64_t particle_id, int64_t at_element,
                                int64_t at_turn );
int64_t oracle_monitor_store_index( const void* monitor, int64_t at_turn,
                                    int64_t particle_id );
void    oracle_cerrf( double x, double y, double* re, double* im );
void    oracle_gauss_field( double x, double y, double sigma_x,
                            double sigma_y, double min_sigma_diff,
                            int skip_Gs, double* Ex, double* Ey,
                            double* Gx, double* Gy );
double  oracle_sin( double x );
double  oracle_cos( double x );
double  oracle_exp( double x );
double  oracle_pow( double x, double y );
int     oracle_uses_libm( void );

#ifdef __cplusplus
}
#endif

#endif /* SIXTRACKLIB_AMD_ORACLE_H__ */

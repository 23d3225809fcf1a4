/* Packed device lattice: the CObjects beam-element buffer flattened into
 * one contiguous stream of 8-byte words, read by the tracking kernels with
 * wave-uniform (scalar) loads.
 *
 * Why not walk the CObjects buffer on the GPU like the reference does
 * (track/track.h:229-582: NS(Object) index -> begin_addr -> struct, with a
 * second pointer chase for Multipole bal)?  Every element would cost two
 * dependent address loads.  Here each element is one self-contained
 * record: header word + parameters inline, so the next record's address is
 * known as soon as the header is read and parameters arrive in the same
 * cache lines.  Constant sub-expressions the reference evaluates per
 * particle are folded at pack time where the folding is bit-identical
 * (e.g. Cavity (PI/180)*lag and ((2 PI)/c)*f are the same IEEE products
 * whether computed on the host or in the kernel).
 *
 * header word: bits 0-7 opcode, bits 8-23 aux (order, flags),
 *              bits 32-63 record length in words including the header.
 */
#ifndef SIXTRACKLIB_AMD_LATTICE_H__
#define SIXTRACKLIB_AMD_LATTICE_H__

#include <stdint.h>

enum st_opcode_e
{
    ST_OP_DRIFT              = 1,   /* [L]                                  */
    ST_OP_DRIFT_EXACT        = 2,   /* [L]                                  */
    ST_OP_MULTIPOLE          = 3,   /* aux: order | 0x8000 curvature;
                                       [bal 0..2n+1] (+ [hxl, hyl, length]) */
    ST_OP_CAVITY             = 4,   /* [V, (pi/180)*lag, ((2pi)/c)*f]       */
    ST_OP_XYSHIFT            = 5,   /* [-dx, -dy]                           */
    ST_OP_SROTATION          = 6,   /* [cos, sin]                           */
    ST_OP_LIMIT_RECT         = 7,   /* [min_x, max_x, min_y, max_y]         */
    ST_OP_LIMIT_ELLIPSE      = 8,   /* [a^2, b^2, a^2 b^2]                  */
    ST_OP_LIMIT_RECT_ELLIPSE = 9,   /* [max_x, max_y, a^2, b^2, a^2 b^2]    */
    ST_OP_DIPEDGE            = 10,  /* [r21, r43]                           */
    ST_OP_MONITOR            = 11,  /* [num_stores, start, skip, out_dev,
                                        max_pid, min_pid, rolling, turn_ord] */
    ST_OP_BB4D               = 12,  /* [11 words of NS(BB4D_data)]          */
    ST_OP_BB6D               = 13,  /* [NS(BB6D_data) words, slices inline] */
    ST_OP_SC_COASTING        = 14,  /* [9 words]                            */
    ST_OP_SC_QGAUSS          = 15,  /* [11 words]                           */
    ST_OP_SC_INTERPOL        = 16,  /* [10 words] + [method, n, z0, dz,
                                        values[n], derivatives[n]]           */
    ST_OP_RF_MULTIPOLE       = 17,  /* aux: order; [V, f, lag, bal, phase]  */
    ST_OP_FAIL               = 18,  /* disabled map: status -1, not lost     */
    ST_OP_UNKNOWN            = 19   /* unknown type: lost, status -1        */
};

#define ST_REC_OP( h )    ( ( unsigned )( ( h ) & 0xffu ) )
#define ST_REC_AUX( h )   ( ( unsigned )( ( ( h ) >> 8 ) & 0xffffu ) )
#define ST_REC_LEN( h )   ( ( unsigned )( ( h ) >> 32 ) )
#define ST_MAKE_HDR( op, aux, len ) ( ( uint64_t )( op ) | \
    ( ( uint64_t )( ( aux ) & 0xffffu ) << 8 ) | ( ( uint64_t )( len ) << 32 ) )
#define ST_MP_CURV_FLAG 0x8000u

/* Device-side particle set: count + 17 FP64 + 4 int64 field pointers, the
 * NS(Particles) layout (sixtracklib/common/particles.h:38-67). */
typedef struct st_dev_particles_s
{
    int64_t n;
    double* q0; double* mass0; double* beta0; double* gamma0; double* p0c;
    double* s; double* x; double* y; double* px; double* py; double* zeta;
    double* psigma; double* delta; double* rpp; double* rvv; double* chi;
    double* charge_ratio;
    int64_t* particle_id; int64_t* at_element; int64_t* at_turn; int64_t* state;
}
st_dev_particles_t;

/* NS(ElemByElemConfig) (output/elem_by_elem_config.h:39-53) */
typedef struct st_ebe_config_s
{
    int64_t order, num_particles_to_store, num_elements_to_store,
            num_turns_to_store, min_particle_id, min_element_id, min_turn,
            max_particle_id, max_element_id, max_turn, is_rolling;
    uint64_t out_store_addr;
}
st_ebe_config_t;

#endif /* SIXTRACKLIB_AMD_LATTICE_H__ */

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
 * record:  word 0  header: bits 0-7 opcode, bits 8-23 aux (order, flags),
 *                   bits 32-63 record length in bytes (header words included)
 *          word 1  element index (position of the object in the buffer)
 *          word 2… payload
 * The stream ends with an ST_OP_END record (followed by zero padding).
 *
 * Drift and DriftExact records carry, after their length, the path length
 * s after the drift on a turn that starts at s = 0: the running sum of the
 * drift lengths from the lattice start, added in element order as the
 * reference adds them to s (track.h:629-660).  A particle that starts its
 * turn at s = +0 has exactly that s after each drift, so the lean loop need
 * not add the length to every particle (it still does for a turn that
 * starts elsewhere).
 *
 * Opcodes with bit 7 clear are the two maps that make up nearly all of a
 * ring (Drift, Multipole): the lean kernel's inner loop tests that one bit
 * to stay on them, and a second bit to tell them apart.  Every other map
 * has bit 7 set and is handled one record at a time.  Lengths are kept in
 * bytes so that advancing to the next record is a single scalar add, and
 * the element index is stored so that the loop needs no element counter
 * (it is read only when a particle is lost).
 */
#ifndef SIXTRACKLIB_AMD_LATTICE_H__
#define SIXTRACKLIB_AMD_LATTICE_H__

#include <stdint.h>

enum st_opcode_e
{
    ST_OP_DRIFT              = 0x01, /* [L, s after the drift]               */
    ST_OP_MULTIPOLE          = 0x02, /* aux: order | 0x8000 curvature;
                                        [bal 0..2n+1] (+ [hxl, hyl, length]) */
    ST_OP_DRIFT_EXACT        = 0x81, /* [L, s after the drift]               */
    ST_OP_CAVITY             = 0x82, /* [V, (pi/180)*lag, ((2pi)/c)*f]       */
    ST_OP_XYSHIFT            = 0x83, /* [-dx, -dy]                           */
    ST_OP_SROTATION          = 0x84, /* [cos, sin]                           */
    ST_OP_LIMIT_RECT         = 0x85, /* [min_x, max_x, min_y, max_y]         */
    ST_OP_LIMIT_ELLIPSE      = 0x86, /* [a^2, b^2, a^2 b^2]                  */
    ST_OP_LIMIT_RECT_ELLIPSE = 0x87, /* [max_x, max_y, a^2, b^2, a^2 b^2]    */
    ST_OP_DIPEDGE            = 0x88, /* [r21, r43]                           */
    ST_OP_MONITOR            = 0x89, /* [num_stores, start, skip, out_dev,
                                         max_pid, min_pid, rolling, turn_ord] */
    ST_OP_BB4D               = 0x8a, /* [11 words of NS(BB4D_data)]          */
    ST_OP_BB6D               = 0x8b, /* [NS(BB6D_data) words, slices inline] */
    ST_OP_SC_COASTING        = 0x8c, /* [9 words]                            */
    ST_OP_SC_QGAUSS          = 0x8d, /* [11 words]                           */
    ST_OP_SC_INTERPOL        = 0x8e, /* [10 words] + [method, n, z0, dz,
                                         values[n], derivatives[n]]           */
    ST_OP_RF_MULTIPOLE       = 0x8f, /* aux: order; [V, f, lag, bal, phase]  */
    ST_OP_FAIL               = 0x90, /* disabled map: status -1, not lost     */
    ST_OP_UNKNOWN            = 0x91, /* unknown type: lost, status -1        */
    ST_OP_END                = 0xff  /* end of the stream                    */
};

#define ST_OP_RARE_BIT     0x80u     /* set for every opcode but the two hot ones */
#define ST_OP_MP_BIT       0x02u     /* Multipole among the hot opcodes           */
#define ST_REC_PAYLOAD     2         /* payload offset in words                   */
#define ST_REC_OP( h )     ( ( unsigned )( ( h ) & 0xffu ) )
#define ST_REC_AUX( h )    ( ( unsigned )( ( ( h ) >> 8 ) & 0xffffu ) )
#define ST_REC_BYTES( h )  ( ( uint32_t )( ( h ) >> 32 ) )
#define ST_REC_LEN( h )    ( ST_REC_BYTES( h ) >> 3 )
#define ST_MAKE_HDR( op, aux, nwords ) ( ( uint64_t )( op ) | \
    ( ( uint64_t )( ( aux ) & 0xffffu ) << 8 ) | ( ( uint64_t )( nwords ) << 35 ) )
/* opcodes the lean kernel runs (no output, no beam-beam / space charge) */
#define ST_OP_IS_LEAN( op ) ( ( op ) == ST_OP_DRIFT || ( op ) == ST_OP_MULTIPOLE || \
    ( ( op ) >= ST_OP_DRIFT_EXACT && ( op ) <= ST_OP_DIPEDGE ) || ( op ) == ST_OP_END )
#define ST_MP_CURV_FLAG 0x8000u
/* Lattice staging through LDS (track_lean_lds.h) copies the stream in
 * pieces of ST_LAT_STAGE_WORDS words; every packed stream is followed by
 * ST_LAT_PAD_WORDS zero words, so that such a piece starting at any record
 * and the kernels' read-ahead of the next record stay inside it. */
#define ST_LAT_STAGE_WORDS 2048
#define ST_LAT_PAD_WORDS   ( 32 + ST_LAT_STAGE_WORDS )

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

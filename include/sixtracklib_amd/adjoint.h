/* sixtracklib_amd/adjoint.h -- differentiable tracking (reverse mode).
 *
 * No counterpart in the reference, which has no derivatives.  Given n
 * particles with initial coordinates (x, px, y, py, zeta, delta), a packed
 * lattice (lattice_pack.h) of Drift, Multipole (no curvature), Cavity,
 * XYShift, SRotation and aperture records, and a target orbit per turn,
 *
 *     loss = sum_i sum_t sum_c  w_c (X_ic(t) - target_c(t))^2
 *
 * where X_i(t) is particle i at the end of turn t.  The forward pass uses
 * the tracker's own element maps (same results as NS(TrackJob_track_until)
 * for particles whose rpp, rvv, psigma follow from delta), stores the lane
 * state every K elements of every turn, and returns the per-particle loss.
 * The reverse pass, one call per turn from the last to the first,
 * recomputes each segment from its checkpoint and applies the transposed
 * Jacobian of every map, giving d loss / d (initial coordinates) and, per
 * multipole, the per-particle data from which
 * NS(Adjoint_reduce_coefficients) forms d loss / d bal (bal[2n] = knl[n]/n!,
 * bal[2n+1] = ksl[n]/n!).  Apertures are not differentiated: a particle an
 * aperture would remove makes the forward pass return status -1.
 *
 * All pointers are device pointers; the caller owns every buffer (sizes
 * below), ``stream`` is a hipStream_t (NULL: the default stream).  Calls
 * are asynchronous on that stream.  Layouts (structure of arrays):
 *   coords, cot   [6][n]
 *   loss          [n]
 *   target        [n_turns][6]
 *   ckpt          NS(Adjoint_checkpoint_doubles)(p) doubles
 *   segbuf        NS(Adjoint_segment_doubles)(p) doubles
 *   coef          [n_multipoles][4][n]
 *   grad_bal      [n_multipoles][stride], accumulated (+=)
 */
#ifndef SIXTRACKLIB_AMD_ADJOINT_H__
#define SIXTRACKLIB_AMD_ADJOINT_H__

#include "sixtracklib_amd/definitions.h"

#if defined( __cplusplus )
extern "C" {
#endif

typedef struct st_adj_problem_s
{
    const uint64_t* lat;        /* packed lattice of n_elem records          */
    const uint32_t* rec_off;    /* [n_elem] word offset of each record       */
    const int32_t* mp_index;    /* [n_elem] multipole slot, -1 for others    */
    int64_t n_elem;             /* elements per turn                         */
    int64_t K;                  /* checkpoint interval (elements)            */
    int64_t n_turns;
    int64_t n;                  /* particles                                 */
    double q0, mass0, beta0, p0c, chi, charge_ratio;   /* one species       */
    double weights[ 6 ];        /* w_c of the loss                           */
    const double* target;       /* [n_turns][6]                              */
    double* ckpt;               /* checkpoints                               */
}
st_adj_problem_t;

ST_API int64_t NS(Adjoint_checkpoint_doubles)( const st_adj_problem_t* p );
ST_API int64_t NS(Adjoint_segment_doubles)( const st_adj_problem_t* p );

/* forward pass; status (device int32) gets -1 OR'ed in when a particle
 * meets an aperture.  0 on a successful launch. */
ST_API int NS(Adjoint_forward)( const st_adj_problem_t* p, int fma, const double* coords,
    double* loss, int32_t* status, void* stream );

/* reverse pass of one turn (call for turn = n_turns - 1 down to 0): cot is
 * the state cotangent at the end of the turn on entry and at its start on
 * return; grad_out (device scalar) multiplies the loss terms. */
ST_API int NS(Adjoint_backward_turn)( const st_adj_problem_t* p, int fma, int64_t turn,
    const double* grad_out, double* segbuf, double* coef, double* cot, void* stream );

/* grad_bal[m][k] += d loss_turn / d bal_m[k] from the coef data of the
 * last reverse pass (mp_order: [n_mp] device int32, orders <= 10) */
ST_API int NS(Adjoint_reduce_coefficients)( const st_adj_problem_t* p, int fma,
    const int32_t* mp_order, int64_t n_mp, const double* coef, double* grad_bal,
    int64_t stride, void* stream );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_ADJOINT_H__ */

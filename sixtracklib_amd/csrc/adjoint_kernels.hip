// Differentiable tracking: forward pass with checkpoints and per-turn
// reverse passes (gfx950; compiled twice like track_kernels.hip).
//
// No counterpart in the reference.  The forward pass runs the same element
// maps as the tracker (track_device.h) on one particle per lane; the loss
// is sum over turns t and coordinates c of w_c (X_c(t) - target_c(t))^2,
// X(t) the state at the end of turn t.  Every K elements of every turn the
// lane state is stored (checkpoint).  The reverse pass for turn t walks the
// turn's segments last to first: it recomputes a segment from its
// checkpoint, storing the state before each element, then applies the
// adjoint maps (adjoint_device.h) in reverse order.  Multipoles write
// (chi px_bar, chi py_bar, x, y) per particle; st_k_adj_reduce turns those
// into the gradient of every bal coefficient (one block per multipole).
//
// Memory layouts (n particles, structure of arrays so a wave's accesses are
// contiguous):
//   coords  [6][n]             x, px, y, py, zeta, delta
//   ckpt    [T][S][9][n]       lane state: the six above + rpp, rvv, psigma
//   segbuf  [K][9][n]          states before each element of one segment
//   coef    [M][4][n]          per multipole: chi px_bar, chi py_bar, x, y
//   cot     [6][n]             cotangent of the state (in / out per turn)
#include <hip/hip_runtime.h>

#include "track_device.h"
#include "adjoint_device.h"
#include "adjoint_launch.h"

#ifndef ST_KSUFFIX
#error "ST_KSUFFIX (kernel name suffix of this build variant) must be defined"
#endif
#define ST_CAT2( a, b ) a##b
#define ST_CAT( a, b ) ST_CAT2( a, b )
#define KN( n ) ST_CAT( n, ST_KSUFFIX )

using namespace st;

namespace
{
typedef __attribute__( ( address_space( 4 ) ) ) const uint32_t cu32_t;
typedef __attribute__( ( address_space( 4 ) ) ) const int32_t ci32_t;

struct AdjArgs
{
    cptr_t lat;
    cu32_t* rec_off;         // word offset of element e's record
    int64_t n_elem, K, n_seg, n_turns, n;
    LaneConst C;
    double w[ 6 ];
    const double* target;    // [T][6]
    const double* coords;
    double* ckpt;
    double* loss;
    int32_t* status;
    int64_t turn;
    const double* grad_out;  // device scalar
    double* segbuf;
    ci32_t* mp_index;        // element -> multipole slot (-1: none)
    double* coef;
    double* cot;
};

ST_DEV void put_lane( double* b, int64_t n, int64_t i, const Lane& L )
{
    b[ 0 * n + i ] = L.x;    b[ 1 * n + i ] = L.px;   b[ 2 * n + i ] = L.y;
    b[ 3 * n + i ] = L.py;   b[ 4 * n + i ] = L.zeta; b[ 5 * n + i ] = L.delta;
    b[ 6 * n + i ] = L.rpp;  b[ 7 * n + i ] = L.rvv;  b[ 8 * n + i ] = L.psigma;
}

ST_DEV void get_lane( const double* b, int64_t n, int64_t i, Lane& L )
{
    L.x = b[ 0 * n + i ];    L.px = b[ 1 * n + i ];   L.y = b[ 2 * n + i ];
    L.py = b[ 3 * n + i ];   L.zeta = b[ 4 * n + i ]; L.delta = b[ 5 * n + i ];
    L.rpp = b[ 6 * n + i ];  L.rvv = b[ 7 * n + i ];  L.psigma = b[ 8 * n + i ];
    L.s = 0.0;
    L.state = 1;
}

// forward map of one record; returns the aperture predicate
ST_DEV bool apply( cptr_t r, Lane& L, const LaneConst& C )
{
    uint64_t const h = r[ 0 ];
    cptr_t const e = r + ST_REC_PAYLOAD;
    switch( ST_REC_OP( h ) )
    {
    case ST_OP_DRIFT: drift( L, rd( e, 0 ) ); return limit_global_ok( L );
    case ST_OP_MULTIPOLE: multipole( L, C.chi, e, ST_REC_AUX( h ), L.delta ); return true;
    case ST_OP_CAVITY: cavity( L, C, e ); return true;
    case ST_OP_XYSHIFT: xyshift( L, e ); return true;
    case ST_OP_SROTATION: srotation( L, e ); return true;
    case ST_OP_LIMIT_RECT: return limit_rect_ok( L, e );
    case ST_OP_LIMIT_ELLIPSE: return limit_ellipse_ok( L, e );
    case ST_OP_LIMIT_RECT_ELLIPSE: return limit_rect_ellipse_ok( L, e );
    default: return false;                   // excluded when the lattice is prepared
    }
}
}  // namespace

__global__ void __launch_bounds__( 256 ) KN( st_k_adj_forward )( AdjArgs a )
{
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    if( i >= a.n ) return;
    int64_t const n = a.n;
    Lane L;
    L.x = a.coords[ 0 * n + i ]; L.px = a.coords[ 1 * n + i ];
    L.y = a.coords[ 2 * n + i ]; L.py = a.coords[ 3 * n + i ];
    L.zeta = a.coords[ 4 * n + i ];
    update_delta( L, a.C, a.coords[ 5 * n + i ] );
    L.s = 0.0;
    L.state = 1;
    double loss = 0.0;
    bool ok = true;
    for( int64_t t = 0; t < a.n_turns; ++t )
    {
        for( int64_t e = 0; e < a.n_elem; ++e )
        {
            if( e % a.K == 0 ) put_lane( a.ckpt + ( ( t * a.n_seg + e / a.K ) * 9 ) * n, n, i, L );
            ok = apply( a.lat + a.rec_off[ e ], L, a.C ) && ok;
        }
        double const X[ 6 ] = { L.x, L.px, L.y, L.py, L.zeta, L.delta };
        for( int c = 0; c < 6; ++c )
        {
            double const d = X[ c ] - a.target[ t * 6 + c ];
            loss += a.w[ c ] * d * d;
        }
    }
    a.loss[ i ] = loss;
    if( !ok ) atomicOr( a.status, -1 );      // an aperture would have removed the particle
}

__global__ void __launch_bounds__( 256 ) KN( st_k_adj_backward_turn )( AdjArgs a )
{
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    if( i >= a.n ) return;
    int64_t const n = a.n, t = a.turn;
    double const g = *a.grad_out;
    Cot c;
    c.x = a.cot[ 0 * n + i ]; c.px = a.cot[ 1 * n + i ]; c.y = a.cot[ 2 * n + i ];
    c.py = a.cot[ 3 * n + i ]; c.zeta = a.cot[ 4 * n + i ]; c.delta = a.cot[ 5 * n + i ];
    for( int64_t j = a.n_seg - 1; j >= 0; --j )
    {
        Lane L;
        get_lane( a.ckpt + ( ( t * a.n_seg + j ) * 9 ) * n, n, i, L );
        int64_t const e0 = j * a.K;
        int64_t const e1 = ( e0 + a.K < a.n_elem ) ? e0 + a.K : a.n_elem;
        for( int64_t e = e0; e < e1; ++e )
        {
            put_lane( a.segbuf + ( ( e - e0 ) * 9 ) * n, n, i, L );
            ( void )apply( a.lat + a.rec_off[ e ], L, a.C );
        }
        if( j == a.n_seg - 1 )
        {
            // loss term of this turn: d/dX of w_c (X_c - target_c)^2
            c.x += g * 2.0 * a.w[ 0 ] * ( L.x - a.target[ t * 6 + 0 ] );
            c.px += g * 2.0 * a.w[ 1 ] * ( L.px - a.target[ t * 6 + 1 ] );
            c.y += g * 2.0 * a.w[ 2 ] * ( L.y - a.target[ t * 6 + 2 ] );
            c.py += g * 2.0 * a.w[ 3 ] * ( L.py - a.target[ t * 6 + 3 ] );
            c.zeta += g * 2.0 * a.w[ 4 ] * ( L.zeta - a.target[ t * 6 + 4 ] );
            c.delta += g * 2.0 * a.w[ 5 ] * ( L.delta - a.target[ t * 6 + 5 ] );
        }
        for( int64_t e = e1 - 1; e >= e0; --e )
        {
            cptr_t const r = a.lat + a.rec_off[ e ];
            uint64_t const h = r[ 0 ];
            cptr_t const prm = r + ST_REC_PAYLOAD;
            Lane P;
            get_lane( a.segbuf + ( ( e - e0 ) * 9 ) * n, n, i, P );
            switch( ST_REC_OP( h ) )
            {
            case ST_OP_DRIFT: drift_adj( P, rd( prm, 0 ), a.C.beta0, c ); break;
            case ST_OP_MULTIPOLE:
            {
                int64_t const m = a.mp_index[ e ];
                double* const cb = a.coef + ( m * 4 ) * n;
                cb[ 0 * n + i ] = a.C.chi * c.px;
                cb[ 1 * n + i ] = a.C.chi * c.py;
                cb[ 2 * n + i ] = P.x;
                cb[ 3 * n + i ] = P.y;
                multipole_adj( P, a.C.chi, prm, ( int )( ST_REC_AUX( h ) & 0x7fffu ), c );
                break;
            }
            case ST_OP_CAVITY: cavity_adj( P, a.C, prm, c ); break;
            case ST_OP_SROTATION: srotation_adj( prm, c ); break;
            default: break;                      // XYShift and apertures: identity
            }
        }
    }
    a.cot[ 0 * n + i ] = c.x; a.cot[ 1 * n + i ] = c.px; a.cot[ 2 * n + i ] = c.y;
    a.cot[ 3 * n + i ] = c.py; a.cot[ 4 * n + i ] = c.zeta; a.cot[ 5 * n + i ] = c.delta;
}

// One block per multipole m: grad_bal[m][2k] += sum_i (-cpx Re z^k + cpy Im z^k),
// grad_bal[m][2k+1] += sum_i (cpx Im z^k + cpy Re z^k), z = x + i y.
__global__ void __launch_bounds__( 256 ) KN( st_k_adj_reduce )( int64_t n,
    const int32_t* mp_order, const double* coef, double* grad_bal, int64_t stride )
{
    __shared__ double red[ 256 ];
    int64_t const m = blockIdx.x;
    int const order = mp_order[ m ];
    double acc[ 2 * ST_ADJ_MAX_ORDER + 2 ];
    for( int k = 0; k < 2 * ST_ADJ_MAX_ORDER + 2; ++k ) acc[ k ] = 0.0;
    const double* const cb = coef + ( m * 4 ) * n;
    for( int64_t i = threadIdx.x; i < n; i += blockDim.x )
    {
        double const cpx = cb[ 0 * n + i ], cpy = cb[ 1 * n + i ];
        double const x = cb[ 2 * n + i ], y = cb[ 3 * n + i ];
        double wr = 1.0, wi = 0.0;
#pragma unroll
        for( int k = 0; k <= ST_ADJ_MAX_ORDER; ++k )
        {
            if( k > order ) break;
            acc[ 2 * k ] += -cpx * wr + cpy * wi;
            acc[ 2 * k + 1 ] += cpx * wi + cpy * wr;
            double const nr = wr * x - wi * y;
            wi = wr * y + wi * x;
            wr = nr;
        }
    }
    for( int k = 0; k < 2 * order + 2; ++k )
    {
        red[ threadIdx.x ] = acc[ k ];
        __syncthreads();
        for( int s = blockDim.x / 2; s > 0; s >>= 1 )
        {
            if( ( int )threadIdx.x < s ) red[ threadIdx.x ] += red[ threadIdx.x + s ];
            __syncthreads();
        }
        if( threadIdx.x == 0 ) grad_bal[ m * stride + k ] += red[ 0 ];
        __syncthreads();
    }
}

// ------------------------------------------------------------ launchers

namespace
{
AdjArgs make_args( const st_adj_problem_t* p )
{
    AdjArgs a{};
    a.lat = ( cptr_t )p->lat;
    a.rec_off = ( cu32_t* )p->rec_off;
    a.n_elem = p->n_elem; a.K = p->K; a.n_seg = ( p->n_elem + p->K - 1 ) / p->K;
    a.n_turns = p->n_turns; a.n = p->n;
    a.C.q0 = p->q0; a.C.mass0 = p->mass0; a.C.beta0 = p->beta0; a.C.p0c = p->p0c;
    a.C.chi = p->chi; a.C.charge_ratio = p->charge_ratio;
    for( int c = 0; c < 6; ++c ) a.w[ c ] = p->weights[ c ];
    a.target = p->target;
    a.ckpt = p->ckpt;
    a.mp_index = ( ci32_t* )p->mp_index;
    return a;
}
}  // namespace

extern "C" hipError_t KN( st_launch_adj_forward )( const st_adj_problem_t* p, const double* coords,
    double* loss, int32_t* status, hipStream_t stream )
{
    if( p->n <= 0 ) return hipSuccess;
    AdjArgs a = make_args( p );
    a.coords = coords; a.loss = loss; a.status = status;
    hipLaunchKernelGGL( KN( st_k_adj_forward ), dim3( ( unsigned )( ( p->n + 255 ) / 256 ) ), dim3( 256 ),
                        0, stream, a );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_adj_backward_turn )( const st_adj_problem_t* p, int64_t turn,
    const double* grad_out, double* segbuf, double* coef, double* cot, hipStream_t stream )
{
    if( p->n <= 0 ) return hipSuccess;
    AdjArgs a = make_args( p );
    a.turn = turn; a.grad_out = grad_out; a.segbuf = segbuf; a.coef = coef; a.cot = cot;
    hipLaunchKernelGGL( KN( st_k_adj_backward_turn ), dim3( ( unsigned )( ( p->n + 255 ) / 256 ) ),
                        dim3( 256 ), 0, stream, a );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_adj_reduce )( const st_adj_problem_t* p, const int32_t* mp_order,
    int64_t n_mp, const double* coef, double* grad_bal, int64_t stride, hipStream_t stream )
{
    if( n_mp <= 0 || p->n <= 0 ) return hipSuccess;
    hipLaunchKernelGGL( KN( st_k_adj_reduce ), dim3( ( unsigned )n_mp ), dim3( 256 ), 0, stream,
                        p->n, mp_order, coef, grad_bal, stride );
    return hipGetLastError();
}

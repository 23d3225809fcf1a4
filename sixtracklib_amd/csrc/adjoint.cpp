// C API of the differentiable tracker (include/sixtracklib_amd/adjoint.h):
// argument checks and dispatch to the _exact / _fma kernel builds.
#include <hip/hip_runtime.h>

#include <string>

#include "sixtracklib_amd/adjoint.h"
#include "adjoint_launch.h"

extern "C" void st_internal_set_error( char const* msg );

namespace
{
bool valid( const st_adj_problem_t* p )
{
    if( p == nullptr || p->lat == nullptr || p->rec_off == nullptr || p->mp_index == nullptr ||
        p->target == nullptr || p->ckpt == nullptr || p->n_elem <= 0 || p->K <= 0 ||
        p->n_turns <= 0 || p->n < 0 || p->n >= ( int64_t )1 << 31 )
    {
        st_internal_set_error( "adjoint: invalid problem description" );
        return false;
    }
    return true;
}

int done( hipError_t e, const char* what )
{
    if( e == hipSuccess ) return 0;
    std::string m = std::string( what ) + ": " + hipGetErrorString( e );
    st_internal_set_error( m.c_str() );
    return -1;
}
}  // namespace

extern "C" {

int64_t NS(Adjoint_checkpoint_doubles)( const st_adj_problem_t* p )
{
    if( p == nullptr || p->K <= 0 ) return -1;
    int64_t const n_seg = ( p->n_elem + p->K - 1 ) / p->K;
    return p->n_turns * n_seg * 9 * p->n;
}

int64_t NS(Adjoint_segment_doubles)( const st_adj_problem_t* p )
{
    if( p == nullptr || p->K <= 0 ) return -1;
    return p->K * 9 * p->n;
}

int NS(Adjoint_forward)( const st_adj_problem_t* p, int fma, const double* coords, double* loss,
    int32_t* status, void* stream )
{
    if( !valid( p ) || coords == nullptr || loss == nullptr || status == nullptr ) return -1;
    hipStream_t const s = static_cast< hipStream_t >( stream );
    return done( fma ? st_launch_adj_forward_fma( p, coords, loss, status, s )
                     : st_launch_adj_forward_exact( p, coords, loss, status, s ), "adjoint forward" );
}

int NS(Adjoint_backward_turn)( const st_adj_problem_t* p, int fma, int64_t turn,
    const double* grad_out, double* segbuf, double* coef, double* cot, void* stream )
{
    if( !valid( p ) || turn < 0 || turn >= p->n_turns || grad_out == nullptr ||
        segbuf == nullptr || coef == nullptr || cot == nullptr )
    {
        st_internal_set_error( "adjoint backward: invalid arguments" );
        return -1;
    }
    hipStream_t const s = static_cast< hipStream_t >( stream );
    return done( fma ? st_launch_adj_backward_turn_fma( p, turn, grad_out, segbuf, coef, cot, s )
                     : st_launch_adj_backward_turn_exact( p, turn, grad_out, segbuf, coef, cot, s ),
                 "adjoint backward" );
}

int NS(Adjoint_reduce_coefficients)( const st_adj_problem_t* p, int fma, const int32_t* mp_order,
    int64_t n_mp, const double* coef, double* grad_bal, int64_t stride, void* stream )
{
    if( !valid( p ) || mp_order == nullptr || coef == nullptr || grad_bal == nullptr ||
        stride < 2 * ST_ADJ_MAX_ORDER + 2 )
    {
        st_internal_set_error( "adjoint reduce: invalid arguments (stride >= 22)" );
        return -1;
    }
    hipStream_t const s = static_cast< hipStream_t >( stream );
    return done( fma ? st_launch_adj_reduce_fma( p, mp_order, n_mp, coef, grad_bal, stride, s )
                     : st_launch_adj_reduce_exact( p, mp_order, n_mp, coef, grad_bal, stride, s ),
                 "adjoint reduce" );
}

}  // extern "C"

// Differentiable tracking: problem descriptor and kernel launchers
// (library-internal; the C API is include/sixtracklib_amd/adjoint.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sixtracklib_amd/adjoint.h"

// largest multipole order the coefficient reduction handles
#define ST_ADJ_MAX_ORDER 10

#define ST_DECLARE_ADJ_LAUNCHERS( SUF )                                                   \
    extern "C" hipError_t st_launch_adj_forward##SUF( const st_adj_problem_t* p,         \
        const double* coords, double* loss, int32_t* status, hipStream_t stream );        \
    extern "C" hipError_t st_launch_adj_backward_turn##SUF( const st_adj_problem_t* p,   \
        int64_t turn, const double* grad_out, double* segbuf, double* coef, double* cot, \
        hipStream_t stream );                                                             \
    extern "C" hipError_t st_launch_adj_reduce##SUF( const st_adj_problem_t* p,          \
        const int32_t* mp_order, int64_t n_mp, const double* coef, double* grad_bal,      \
        int64_t stride, hipStream_t stream );

ST_DECLARE_ADJ_LAUNCHERS( _exact )
ST_DECLARE_ADJ_LAUNCHERS( _fma )

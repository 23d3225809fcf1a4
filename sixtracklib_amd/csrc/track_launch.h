// Host-side prototypes of the kernel launchers (two build variants of
// track_kernels.hip: _exact without FP contraction, _fma with it).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

#define ST_DECLARE_LAUNCHERS( SUF )                                              \
    extern "C" hipError_t st_launch_track_until##SUF( const st_dev_particles_t* P, \
        const uint64_t* lat, int64_t n_elem, int64_t until_turn, int full,       \
        int32_t* status, int block, hipStream_t stream );                         \
    extern "C" hipError_t st_launch_track_ebe##SUF( const st_dev_particles_t* P,  \
        const uint64_t* lat, int64_t n_elem, int64_t until_turn,                  \
        const st_ebe_config_t* cfg, int64_t cap, const st_dev_particles_t* O,       \
        int32_t* status, int64_t from_turn,                                        \
        int block, hipStream_t stream );                                          \
    extern "C" hipError_t st_launch_track_line##SUF( const st_dev_particles_t* P, \
        const uint64_t* lat, int64_t n_line, int finish_turn, int32_t* status,    \
        int block, hipStream_t stream );

ST_DECLARE_LAUNCHERS( _exact )
ST_DECLARE_LAUNCHERS( _fma )

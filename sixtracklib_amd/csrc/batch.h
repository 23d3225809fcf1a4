// Batched multi-seed tracking: device-side descriptors shared by the host
// driver (batch.cpp) and the kernels (batch_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

// One machine seed: its particle set, its packed lattice, its active list.
typedef struct st_batch_seed_s
{
    st_dev_particles_t P;       // particle set (device view)
    const uint64_t* lat;        // packed lean lattice (device), END-terminated
    int64_t n_elem;             // elements per turn
    int32_t* list;              // active-particle list, capacity P.n
}
st_batch_seed_t;

// Work distribution of one launch: seeds are split into queues (one per
// XCD when grouped), a queue is drained tile by tile (64 particles).
typedef struct st_batch_sched_s
{
    const st_batch_seed_t* seeds;   // [n_seeds]
    int32_t* counts;                // [2 * n_seeds]: active count, smallest at_turn
    const int32_t* queue_seeds;     // seeds of queue q: [queue_begin[q], queue_begin[q+1])
    const int32_t* queue_begin;     // [n_queues + 1]
    int32_t* queue_next;            // [n_queues] next tile to hand out
    // Balanced form (split != nullptr): each seed's list holds the
    // particles at the smallest at_turn of all seeds (*gmin) at its front,
    // split[2k] of them, and the others at its back, split[2k + 1] of them
    // ending at index P.n - 1; a wave tracks at most one tile per launch,
    // front tiles of every seed before back tiles.
    int32_t* split;                 // [2 * n_seeds] or nullptr
    int32_t* gmin;                  // [1] smallest at_turn with turns to go
    unsigned long long* progress;   // [1] turns done by the waves of a balanced launch
    int32_t n_seeds;
    int32_t n_queues;
    int32_t wave_prio;              // waves prioritised by turns left (track_lean.h)
}
st_batch_sched_t;

#define ST_DECLARE_BATCH_LAUNCHERS( SUF )                                            \
    extern "C" hipError_t st_launch_batch_reset##SUF( const st_batch_sched_t* s,     \
        hipStream_t stream );                                                         \
    extern "C" hipError_t st_launch_batch_select##SUF( const st_batch_sched_t* s,    \
        int64_t max_particles, int64_t until_turn, hipStream_t stream );              \
    extern "C" hipError_t st_launch_batch_select_split##SUF( const st_batch_sched_t* s, \
        int64_t max_particles, int64_t until_turn, hipStream_t stream );              \
    extern "C" hipError_t st_launch_batch_track##SUF( const st_batch_sched_t* s,     \
        int64_t until_turn, int64_t max_turns, int32_t* status, int block, int grid, \
        int particles_per_lane, int heavy, int staging, hipStream_t stream );          \
    extern "C" hipError_t st_launch_batch_track_mix##SUF( const st_batch_sched_t* s, \
        int64_t until_turn, int64_t max_turns, int32_t* status, int grid,            \
        int64_t pair_waves, hipStream_t stream );

ST_DECLARE_BATCH_LAUNCHERS( _exact )
ST_DECLARE_BATCH_LAUNCHERS( _fma )

/* sixtracklib_amd/track_batch.h -- many machine seeds in one launch.
 *
 * No counterpart in the reference.  Its nearest relatives are the CPU job's
 * serial loop over particle sets (common/internal/track_job_cpu.cpp:
 * 360-376) and the CUDA job, which asserts a single particle set
 * (cuda/track/track_job.cpp:2347); a user of the reference tracks N seeds
 * with N jobs one after the other.  Here N (lattice, particle-set) pairs
 * are tracked by the same persistent launches: by default every launch
 * gives each resident wave one tile of any seed, the particles furthest
 * behind first (balance=1); with balance=0 the chunks of turns of a
 * track_until call are captured in one HIP graph.  Each seed's result is
 * bit-identical to a plain NS(TrackJob_track_until) on that seed (same
 * per-particle code; fma=1 selects the fused build for both).
 *
 * Lattices must be "lean" (Drift, DriftExact, Multipole, Cavity, XYShift,
 * SRotation, LimitRect / LimitEllipse / LimitRectEllipse, DipoleEdge);
 * the first particle set of each particle buffer is tracked.
 *
 * config_str tokens (space or comma separated):
 *   fma=0|1               arithmetic build (default 0, bit-exact)
 *   block=64|128|256      threads per block of the persistent launch
 *   waves_per_cu=N        resident waves per CU the grid is sized for (1-32)
 *   queues=1|8            8: seeds grouped per XCD (one work queue each,
 *                         lattice shared in that XCD's L2), 1: one queue
 *   turns_per_launch=K    turns per chunk (0: from the work size)
 *   graph=0|1             capture the chunks of a call in a HIP graph (1)
 *   ppl=1|2               particles per lane (1): with 2 every lane tracks
 *                         two particles through the shared record walk
 *                         (128-particle tiles; results identical)
 *   wave_prio=0|1         waves with more turns left in their tile get
 *                         issue priority (1; results identical)
 *   staging=scalar|lds    lattice read through scalar loads by every wave
 *                         (scalar) or copied piece by piece into LDS by
 *                         256-lane workgroups that walk it together (lds;
 *                         one particle per lane; results identical)
 *   balance=0|1           1 (default): launches of one tile per resident
 *                         wave, the particles furthest behind first, the
 *                         host reading the counts back between launch
 *                         groups (no graph); 0: chunks of turns, every wave
 *                         taking tiles until none are left (graph-captured);
 *                         results identical
 * NS(TrackBatch_autotune) picks block, waves_per_cu, queues, particles per
 * lane and lattice staging by timing short trials on the batch itself, the
 * three fastest twice (state restored afterwards).
 */
#ifndef SIXTRACKLIB_AMD_TRACK_BATCH_H__
#define SIXTRACKLIB_AMD_TRACK_BATCH_H__

#include "sixtracklib_amd/definitions.h"
#include "sixtracklib_amd/buffer.h"

#if defined( __cplusplus )
extern "C" {
#endif

typedef struct NS(TrackBatch) NS(TrackBatch);

/* NULL on failure (NS(Hip_last_error_str) says why).  The caller keeps
 * ownership of the buffers; they must stay alive while the batch exists. */
ST_API NS(TrackBatch)* NS(TrackBatch_new)( int64_t n_seeds,
    NS(Buffer)* const* lattices, NS(Buffer)* const* particles, char const* config_str );
ST_API void NS(TrackBatch_delete)( NS(TrackBatch)* batch );

/* Every seed to until_turn; returns the OR of the per-particle statuses
 * (0 success, -1 failure), like NS(TrackJob_track_until). */
ST_API NS(track_status_t) NS(TrackBatch_track_until)( NS(TrackBatch)* batch, int64_t until_turn );

/* device -> caller's particle buffers / caller's buffers -> device */
ST_API int NS(TrackBatch_collect)( NS(TrackBatch)* batch );
ST_API int NS(TrackBatch_push)( NS(TrackBatch)* batch );

ST_API int NS(TrackBatch_set_config)( NS(TrackBatch)* batch, char const* config_str );
/* writes the effective configuration (NUL-terminated); returns its length */
ST_API int64_t NS(TrackBatch_get_config)( const NS(TrackBatch)* batch, char* out, int64_t out_len );
/* times each candidate configuration for trial_turns turns on the batch,
 * after 10 x trial_turns turns tracked with the current configuration (the
 * trials see the batch past its early losses), keeps the fastest and
 * restores the batch state; 0 on success */
ST_API int NS(TrackBatch_autotune)( NS(TrackBatch)* batch, int64_t trial_turns );
/* as NS(TrackBatch_autotune) with the trials starting after skip_turns
 * turns (0: at the current state) */
ST_API int NS(TrackBatch_autotune_window)( NS(TrackBatch)* batch, int64_t skip_turns,
    int64_t trial_turns );

ST_API int64_t NS(TrackBatch_num_seeds)( const NS(TrackBatch)* batch );
ST_API int64_t NS(TrackBatch_total_particles)( const NS(TrackBatch)* batch );
/* device time of the last track_until (HIP events around the graph) */
ST_API double NS(TrackBatch_last_kernel_time_ms)( const NS(TrackBatch)* batch );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_TRACK_BATCH_H__ */

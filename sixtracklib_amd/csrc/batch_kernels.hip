// Batched multi-seed tracking kernels for gfx950 (compiled twice, like
// track_kernels.hip: _exact without FP contraction, _fma with it).
//
// No counterpart in the reference: its CPU job loops over particle sets
// one after the other (internal/track_job_cpu.cpp:360-376) and its CUDA
// job tracks a single set (cuda/track/track_job.cpp:2347).  Every lane
// runs the same per-particle map as the plain lean kernel
// (track_lean_lane), so each seed's results are those of a plain
// track_until on that seed alone.
//
// One chunk of turns is three launches:
//   reset   queue counters and per-seed counts;
//   select  per seed, the particles with turns to go -> the seed's list;
//   track   a persistent grid: every wave takes 64-particle tiles from its
//           queue (one queue per XCD when seeds are grouped, so the waves
//           of an XCD share one or two lattices in its L2), tracks a tile
//           for up to max_turns turns, takes the next; when its queue is
//           empty it takes tiles from the other queues.  Tiles are finite
//           and counters only grow, so every wave reaches the exit.
#include <hip/hip_runtime.h>

#include "track_device.h"
#include "track_lean.h"
#include "track_lean_lds.h"
#include "batch.h"

#ifndef ST_KSUFFIX
#error "ST_KSUFFIX (kernel name suffix of this build variant) must be defined"
#endif
#define ST_CAT2( a, b ) a##b
#define ST_CAT( a, b ) ST_CAT2( a, b )
#define KN( n ) ST_CAT( n, ST_KSUFFIX )

#ifndef ST_LEAN_WAVES_PER_EU
#define ST_LEAN_WAVES_PER_EU 8
#endif

using namespace st;

namespace
{
// read-only per-launch tables through the constant address space: the
// indices are wave-uniform, so these become scalar loads
typedef __attribute__( ( address_space( 4 ) ) ) const int32_t ci32_t;

static_assert( sizeof( st_dev_particles_t ) == 22 * 8, "particle view: count + 21 pointers" );
static_assert( sizeof( st_batch_seed_t ) == 25 * 8, "seed descriptor: view, lat, n_elem, list" );

// Seed descriptor k, read as 25 words through the constant address space.
struct SeedView
{
    st_dev_particles_t P;
    cptr_t lat;
    int64_t n_elem;
    int32_t* list;
};

ST_DEV SeedView load_seed( const st_batch_seed_t* seeds, int k )
{
    cptr_t const w = ( cptr_t )seeds + 25 * k;
    SeedView v;
    v.P.n = ( int64_t )w[ 0 ];
    v.P.q0 = ( double* )w[ 1 ];      v.P.mass0 = ( double* )w[ 2 ];
    v.P.beta0 = ( double* )w[ 3 ];   v.P.gamma0 = ( double* )w[ 4 ];
    v.P.p0c = ( double* )w[ 5 ];     v.P.s = ( double* )w[ 6 ];
    v.P.x = ( double* )w[ 7 ];       v.P.y = ( double* )w[ 8 ];
    v.P.px = ( double* )w[ 9 ];      v.P.py = ( double* )w[ 10 ];
    v.P.zeta = ( double* )w[ 11 ];   v.P.psigma = ( double* )w[ 12 ];
    v.P.delta = ( double* )w[ 13 ];  v.P.rpp = ( double* )w[ 14 ];
    v.P.rvv = ( double* )w[ 15 ];    v.P.chi = ( double* )w[ 16 ];
    v.P.charge_ratio = ( double* )w[ 17 ];
    v.P.particle_id = ( int64_t* )w[ 18 ]; v.P.at_element = ( int64_t* )w[ 19 ];
    v.P.at_turn = ( int64_t* )w[ 20 ];     v.P.state = ( int64_t* )w[ 21 ];
    v.lat = ( cptr_t )w[ 22 ];
    v.n_elem = ( int64_t )w[ 23 ];
    v.list = ( int32_t* )w[ 24 ];
    return v;
}
}  // namespace

__global__ void __launch_bounds__( 256 ) KN( st_k_batch_reset )( st_batch_sched_t s )
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    int const stride = gridDim.x * blockDim.x;
    for( int q = i; q < s.n_queues; q += stride ) s.queue_next[ q ] = 0;
    for( int k = i; k < s.n_seeds; k += stride )
    {
        s.counts[ 2 * k ] = 0;
        s.counts[ 2 * k + 1 ] = INT32_MAX;
        if( s.split != nullptr ) { s.split[ 2 * k ] = 0; s.split[ 2 * k + 1 ] = 0; }
    }
    if( i == 0 && s.gmin != nullptr ) *s.gmin = INT32_MAX;
    if( i == 0 && s.progress != nullptr ) *s.progress = 0ull;
}

__global__ void __launch_bounds__( 256 ) KN( st_k_batch_select )( st_batch_sched_t s,
    int64_t until_turn )
{
    SeedView const S = load_seed( s.seeds, ( int )blockIdx.y );
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    if( ( int64_t )blockIdx.x * blockDim.x >= S.P.n ) return;      // whole block past the set
    select_active_lane( S.P, i, until_turn, S.list, &s.counts[ 2 * blockIdx.y ] );
}

// Balanced form, pass 1: the smallest at_turn (clamped to int32) among the
// particles of all seeds with turns to go -> *s.gmin.  One atomic per block.
__global__ void __launch_bounds__( 256 ) KN( st_k_batch_minturn )( st_batch_sched_t s,
    int64_t until_turn )
{
    cptr_t const w = ( cptr_t )s.seeds + 25 * blockIdx.y;
    int64_t const n = ( int64_t )w[ 0 ];
    if( ( int64_t )blockIdx.x * blockDim.x >= n ) return;            // whole block past the set
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    int64_t const t = ( i < n ) ? ( ( const int64_t* )w[ 20 ] )[ i ] : 0;
    bool const want = i < n && ( ( const int64_t* )w[ 21 ] )[ i ] == 1 && t < until_turn;
    int32_t m = want ? ( int32_t )( t < 0 ? 0 : ( t > INT32_MAX ? INT32_MAX : t ) ) : INT32_MAX;
    for( int off = 32; off > 0; off >>= 1 ) m = min( m, __shfl_xor( m, off ) );
    __shared__ int32_t part[ 4 ];
    if( __lane_id() == 0 ) part[ threadIdx.x >> 6 ] = m;
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        int32_t b = part[ 0 ];
        for( int k = 1; k < ( int )( blockDim.x >> 6 ); ++k ) b = min( b, part[ k ] );
        if( b < INT32_MAX ) atomicMin( s.gmin, b );
    }
}

// Balanced form, pass 2: the particles with turns to go, those at *s.gmin
// to the front of the seed's list (split[2k]), the others to its back
// (split[2k + 1], filled from index P.n - 1 down).  One atomic add per
// block and part; the order within a part is of no consequence.
__global__ void __launch_bounds__( 256 ) KN( st_k_batch_select_split )( st_batch_sched_t s,
    int64_t until_turn )
{
    int const k = ( int )blockIdx.y;
    cptr_t const w = ( cptr_t )s.seeds + 25 * k;
    int64_t const n = ( int64_t )w[ 0 ];
    if( ( int64_t )blockIdx.x * blockDim.x >= n ) return;            // whole block past the set
    int32_t* const list = ( int32_t* )w[ 24 ];
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    int64_t const t = ( i < n ) ? ( ( const int64_t* )w[ 20 ] )[ i ] : 0;
    bool const want = i < n && ( ( const int64_t* )w[ 21 ] )[ i ] == 1 && t < until_turn;
    bool const front = want && t <= ( int64_t )*s.gmin;
    bool const back = want && !front;
    uint64_t const mf = ballot( front ), mb = ballot( back );
    __shared__ int32_t cf[ 4 ], cb[ 4 ], base[ 2 ];
    int const wv = ( int )( threadIdx.x >> 6 );
    if( __lane_id() == 0 ) { cf[ wv ] = __popcll( mf ); cb[ wv ] = __popcll( mb ); }
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        int32_t tf = 0, tb = 0;
        for( int q = 0; q < ( int )( blockDim.x >> 6 ); ++q ) { tf += cf[ q ]; tb += cb[ q ]; }
        base[ 0 ] = tf > 0 ? atomicAdd( &s.split[ 2 * k ], tf ) : 0;
        base[ 1 ] = tb > 0 ? atomicAdd( &s.split[ 2 * k + 1 ], tb ) : 0;
    }
    __syncthreads();
    int32_t of = base[ 0 ], ob = base[ 1 ];
    for( int q = 0; q < wv; ++q ) { of += cf[ q ]; ob += cb[ q ]; }
    uint64_t const below = ( __lane_id() == 0 ) ? 0u : ( ~0ull >> ( 64 - __lane_id() ) );
    if( front ) list[ of + __popcll( mf & below ) ] = ( int32_t )i;
    if( back ) list[ n - 1 - ( ob + __popcll( mb & below ) ) ] = ( int32_t )i;
}

// Issue-priority rule of a track launch (WavePrio, track_lean.h): none
// without wave_prio; balanced launches keep their waves level with the
// mean progress of the min(tiles, resident waves) waves that have a tile;
// the persistent-chunk form bands by turns left.
template< int TILE >
ST_DEV WavePrio launch_priority( const st_batch_sched_t& s, int64_t max_turns )
{
    WavePrio wp;
    if( !s.wave_prio || max_turns <= 0 ) return wp;
    wp.k = max_turns;
    if( s.split == nullptr || s.progress == nullptr ) return wp;
    ci32_t* const split = ( ci32_t* )s.split;
    int64_t tiles = 0;
    for( int k = 0; k < 2 * s.n_seeds; ++k ) tiles += ( split[ k ] + TILE - 1 ) / TILE;
    int64_t const waves = ( int64_t )gridDim.x * ( blockDim.x >> 6 );
    wp.progress = s.progress;
    wp.n_waves = tiles < waves ? tiles : waves;
    return wp;
}

// Tile t of a queue (seeds qseeds[qb..qe)) -> its seed, the part of the
// seed's list it indexes (list, count) and the tile's index in that part.
// False when the queue holds t or fewer tiles.  TILE particles per tile.
template< int TILE >
ST_DEV bool find_tile( const st_batch_sched_t& s, int qb, int qe, int t, int& seed,
                       const int32_t*& list, int32_t& count, int& tile )
{
    ci32_t* const qseeds = ( ci32_t* )s.queue_seeds;
    if( s.split == nullptr )
    {
        ci32_t* const counts = ( ci32_t* )s.counts;
        for( int k = qb; k < qe; ++k )
        {
            int const sd = qseeds[ k ];
            int const n = counts[ 2 * sd ];
            int const nt = ( n + TILE - 1 ) / TILE;
            if( t < nt )
            {
                seed = sd; count = n; tile = t;
                list = ( const int32_t* )( ( cptr_t )s.seeds + 25 * sd )[ 24 ];
                return true;
            }
            t -= nt;
        }
        return false;
    }
    ci32_t* const split = ( ci32_t* )s.split;
    for( int part = 0; part < 2; ++part )
        for( int k = qb; k < qe; ++k )
        {
            int const sd = qseeds[ k ];
            int const n = split[ 2 * sd + part ];
            int const nt = ( n + TILE - 1 ) / TILE;
            if( t < nt )
            {
                cptr_t const w = ( cptr_t )s.seeds + 25 * sd;
                seed = sd; count = n; tile = t;
                list = ( const int32_t* )w[ 24 ] + ( part == 0 ? 0 : ( int64_t )w[ 0 ] - n );
                return true;
            }
            t -= nt;
        }
    return false;
}

// Tile loop of the persistent track kernels (HEAVY: lattices that also
// hold the maps outside the lean set).
template< bool HEAVY >
ST_DEV void batch_track_tiles( const st_batch_sched_t& s, int64_t until_turn, int64_t max_turns,
                               int32_t* status )
{
    ci32_t* const qbegin = ( ci32_t* )s.queue_begin;
    int const nq = s.n_queues;
    int const q0 = ( int )( blockIdx.x % ( unsigned )nq );
    bool const one_tile = s.split != nullptr;
    WavePrio const wp = launch_priority< 64 >( s, max_turns );
    for( int qi = 0; qi < nq; ++qi )
    {
        int const q = ( q0 + qi ) % nq;
        int const qb = qbegin[ q ], qe = qbegin[ q + 1 ];
        for( ;; )
        {
            int t = 0;
            if( __lane_id() == 0 ) t = atomicAdd( &s.queue_next[ q ], 1 );
            t = __builtin_amdgcn_readfirstlane( __shfl( t, 0 ) );
            int seed = 0, tile = 0;
            int32_t count = 0;
            const int32_t* list = nullptr;
            if( !find_tile< 64 >( s, qb, qe, t, seed, list, count, tile ) ) break;   // queue drained
            if constexpr( HEAVY )
            {
                // the particle set is read from the seed descriptor where
                // it is used (its first 22 words are the set's view)
                cptr_t const w = ( cptr_t )s.seeds + 25 * seed;
                track_lean_lane< true >( PDesc{ w }, list, count, ( int64_t )tile * 64 + __lane_id(),
                                         ( cptr_t )w[ 22 ], ( int64_t )w[ 23 ], until_turn, max_turns,
                                         status, wp );
            }
            else
            {
                SeedView const S = load_seed( s.seeds, seed );
                track_lean_lane< false >( PDirect{ S.P }, list, count, ( int64_t )tile * 64 + __lane_id(),
                                          S.lat, S.n_elem, until_turn, max_turns, status, wp );
            }
            if( one_tile ) return;
        }
    }
}

__global__ void __launch_bounds__( 256, ST_LEAN_WAVES_PER_EU )
KN( st_k_batch_track )( st_batch_sched_t s, int64_t until_turn, int64_t max_turns,
                        int32_t* status )
{
    batch_track_tiles< false >( s, until_turn, max_turns, status );
}

// The same for lattices with monitors, beam-beam, space charge or RF
// multipoles: those records run out of line (heavy_step), the Drift /
// Multipole runs keep the lean loop's registers and occupancy.
__global__ void __launch_bounds__( 256, ST_LEAN_WAVES_PER_EU )
KN( st_k_batch_track_heavy )( st_batch_sched_t s, int64_t until_turn, int64_t max_turns,
                              int32_t* status )
{
    batch_track_tiles< true >( s, until_turn, max_turns, status );
}

// Lattice staged through LDS (track_lean_lds.h): a workgroup of 256 lanes
// takes one 256-particle tile of one particle set at a time and walks the
// lattice with its four waves together.  16 KB of LDS per workgroup, 8
// workgroups per CU.
__global__ void __launch_bounds__( 256, ST_LEAN_WAVES_PER_EU )
KN( st_k_batch_track_lds )( st_batch_sched_t s, int64_t until_turn, int64_t max_turns,
                            int32_t* status )
{
    __shared__ uint64_t stage[ ST_LAT_STAGE_WORDS ];
    __shared__ int next_tile;
    ci32_t* const qbegin = ( ci32_t* )s.queue_begin;
    int const nq = s.n_queues;
    int const q0 = ( int )( blockIdx.x % ( unsigned )nq );
    bool const one_tile = s.split != nullptr;
    for( int qi = 0; qi < nq; ++qi )
    {
        int const q = ( q0 + qi ) % nq;
        int const qb = qbegin[ q ], qe = qbegin[ q + 1 ];
        for( ;; )
        {
            if( threadIdx.x == 0 ) next_tile = atomicAdd( &s.queue_next[ q ], 1 );
            __syncthreads();
            int const t = next_tile;
            __syncthreads();                  // read by all before the next write
            int seed = 0, tile = 0;
            int32_t count = 0;
            const int32_t* list = nullptr;
            if( !find_tile< 256 >( s, qb, qe, t, seed, list, count, tile ) ) break;  // queue drained
            SeedView const S = load_seed( s.seeds, seed );
            track_lean_block( S.P, list, count, ( int64_t )tile * 256, S.lat, S.n_elem, until_turn,
                              max_turns, status,
                              ( __attribute__( ( address_space( 3 ) ) ) uint64_t* )stage );
            if( one_tile ) return;
        }
    }
}

// As st_k_batch_track with 128-particle tiles, two particles per lane, at
// the same eight waves per SIMD (track_lean_lane2 keeps the pair in 64
// VGPRs; its rare maps spill a few registers to scratch).  The particle set
// is read from the seed descriptor where it is used, so that no field
// address stays in SGPRs across the element loop.
__global__ void __launch_bounds__( 256, ST_LEAN_WAVES_PER_EU )
KN( st_k_batch_track2 )( st_batch_sched_t s, int64_t until_turn, int64_t max_turns,
                         int32_t* status )
{
    ci32_t* const qbegin = ( ci32_t* )s.queue_begin;
    int const nq = s.n_queues;
    int const q0 = ( int )( blockIdx.x % ( unsigned )nq );
    bool const one_tile = s.split != nullptr;
    WavePrio const wp = launch_priority< 128 >( s, max_turns );
    for( int qi = 0; qi < nq; ++qi )
    {
        int const q = ( q0 + qi ) % nq;
        int const qb = qbegin[ q ], qe = qbegin[ q + 1 ];
        for( ;; )
        {
            int t = 0;
            if( __lane_id() == 0 ) t = atomicAdd( &s.queue_next[ q ], 1 );
            t = __builtin_amdgcn_readfirstlane( __shfl( t, 0 ) );
            int seed = 0, tile = 0;
            int32_t count = 0;
            const int32_t* list = nullptr;
            if( !find_tile< 128 >( s, qb, qe, t, seed, list, count, tile ) ) break;  // queue drained
            cptr_t const w = ( cptr_t )s.seeds + 25 * seed;
            track_lean_lane2( PDesc{ w }, list, count, ( int64_t )tile * 128, ( cptr_t )w[ 22 ],
                              ( int64_t )w[ 23 ], until_turn, max_turns, status, wp );
            if( one_tile ) return;
        }
    }
}

// Mixed launch of one seed whose active particles all sit at one turn (the
// front part of its list, split[0] of them): waves 0 .. pair_waves - 1 take
// 128 particles each, two per lane (track_lean_lane2), the others share the
// rest evenly, one per lane (track_lean_lane).  With a grid of 8 workgroups
// of 256 per CU the dispatcher gives every SIMD one wave of each run of
// n_cu consecutive workgroups, so pair_waves = m * 4 * n_cu puts m pair
// waves and 8 - m single waves on every SIMD: all resident waves busy, the
// same work on every SIMD, where one particle per lane would need a second,
// part-filled launch and two per lane would leave waves idle.  The issue
// priority keeps the waves level in turns (the pairs, slower per turn, go
// first when behind).
__global__ void __launch_bounds__( 256, ST_LEAN_WAVES_PER_EU )
KN( st_k_batch_track_mix )( st_batch_sched_t s, int64_t until_turn, int64_t max_turns,
                            int32_t* status, int64_t pair_waves )
{
    ci32_t* const split = ( ci32_t* )s.split;
    int64_t const n = split[ 0 ];
    int64_t const W = ( int64_t )gridDim.x * ( blockDim.x >> 6 );
    int64_t const w = ( int64_t )blockIdx.x * ( blockDim.x >> 6 ) + ( threadIdx.x >> 6 );
    cptr_t const sd = ( cptr_t )s.seeds;                      // seed 0
    const int32_t* const list = ( const int32_t* )sd[ 24 ];
    cptr_t const lat = ( cptr_t )sd[ 22 ];
    int64_t const n_elem = ( int64_t )sd[ 23 ];
    WavePrio wp;
    if( s.wave_prio && max_turns > 0 && s.progress != nullptr )
    {
        wp.k = max_turns;
        wp.progress = s.progress;
        wp.n_waves = W;
    }
    int64_t const pairs = pair_waves < W ? pair_waves : W;
    if( w < pairs )
    {
        // 128 particles per pair wave; with every wave a pair wave, n is
        // spread evenly over them instead (at most 128 each)
        int64_t begin = 128 * w, end = begin + 128 < n ? begin + 128 : n;
        if( pairs == W )
        {
            int64_t const q = n / W, r = n % W;
            begin = w * q + ( w < r ? w : r );
            end = begin + q + ( w < r ? 1 : 0 );
        }
        track_lean_lane2( PDesc{ sd }, list, end, begin, lat, n_elem, until_turn, max_turns, status, wp );
        return;
    }
    int64_t const U = W - pairs;                               // single waves
    int64_t const rest = n > 128 * pairs ? n - 128 * pairs : 0;
    int64_t const u = w - pairs;
    int64_t const q = rest / U, r = rest % U;
    int64_t const begin = 128 * pairs + u * q + ( u < r ? u : r );
    int64_t const end = begin + q + ( u < r ? 1 : 0 );
    track_lean_lane< false >( PDesc{ sd }, list, end, begin + __lane_id(), lat, n_elem, until_turn,
                              max_turns, status, wp );
}

// ------------------------------------------------------------ launchers

extern "C" hipError_t KN( st_launch_batch_reset )( const st_batch_sched_t* s, hipStream_t stream )
{
    hipLaunchKernelGGL( KN( st_k_batch_reset ), dim3( 1 ), dim3( 256 ), 0, stream, *s );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_batch_select )( const st_batch_sched_t* s,
    int64_t max_particles, int64_t until_turn, hipStream_t stream )
{
    if( max_particles <= 0 || s->n_seeds <= 0 ) return hipSuccess;
    unsigned const gx = ( unsigned )( ( max_particles + 255 ) / 256 );
    hipLaunchKernelGGL( KN( st_k_batch_select ), dim3( gx, ( unsigned )s->n_seeds ), dim3( 256 ),
                        0, stream, *s, until_turn );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_batch_select_split )( const st_batch_sched_t* s,
    int64_t max_particles, int64_t until_turn, hipStream_t stream )
{
    if( max_particles <= 0 || s->n_seeds <= 0 ) return hipSuccess;
    if( s->split == nullptr || s->gmin == nullptr ) return hipErrorInvalidValue;
    unsigned const gx = ( unsigned )( ( max_particles + 255 ) / 256 );
    hipLaunchKernelGGL( KN( st_k_batch_minturn ), dim3( gx, ( unsigned )s->n_seeds ), dim3( 256 ),
                        0, stream, *s, until_turn );
    hipError_t const e = hipGetLastError();
    if( e != hipSuccess ) return e;
    hipLaunchKernelGGL( KN( st_k_batch_select_split ), dim3( gx, ( unsigned )s->n_seeds ),
                        dim3( 256 ), 0, stream, *s, until_turn );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_batch_track_mix )( const st_batch_sched_t* s,
    int64_t until_turn, int64_t max_turns, int32_t* status, int grid, int64_t pair_waves,
    hipStream_t stream )
{
    if( s->n_seeds != 1 || s->split == nullptr || grid <= 0 ) return hipErrorInvalidValue;
    hipLaunchKernelGGL( KN( st_k_batch_track_mix ), dim3( ( unsigned )grid ), dim3( 256 ), 0, stream,
                        *s, until_turn, max_turns, status, pair_waves );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_batch_track )( const st_batch_sched_t* s,
    int64_t until_turn, int64_t max_turns, int32_t* status, int block, int grid,
    int particles_per_lane, int heavy, int staging, hipStream_t stream )
{
    if( s->n_seeds <= 0 || s->n_queues <= 0 ) return hipSuccess;
    if( block <= 0 || block > 256 || ( block % 64 ) != 0 ) block = 256;
    if( grid <= 0 ) grid = 1;
    if( staging && !heavy )
        hipLaunchKernelGGL( KN( st_k_batch_track_lds ), dim3( ( unsigned )grid ), dim3( 256 ), 0,
                            stream, *s, until_turn, max_turns, status );
    else if( heavy )
        hipLaunchKernelGGL( KN( st_k_batch_track_heavy ), dim3( ( unsigned )grid ),
                            dim3( ( unsigned )block ), 0, stream, *s, until_turn, max_turns, status );
    else if( particles_per_lane == 2 )
        hipLaunchKernelGGL( KN( st_k_batch_track2 ), dim3( ( unsigned )grid ),
                            dim3( ( unsigned )block ), 0, stream, *s, until_turn, max_turns, status );
    else
        hipLaunchKernelGGL( KN( st_k_batch_track ), dim3( ( unsigned )grid ),
                            dim3( ( unsigned )block ), 0, stream, *s, until_turn, max_turns, status );
    return hipGetLastError();
}

// Persistent-kernel scheduling of lean until-turn tracking (host side),
// shared by NS(TrackBatch) (many lattices) and NS(TrackJob) (one lattice,
// one or more particle sets).  Library-internal C++.
//
// A "seed" is one (particle set, lean lattice) pair.  run(U, C, K) enqueues
// C chunks, each
//   reset   tile counters and per-seed [count, smallest at_turn];
//   select  per seed, the particles with turns to go -> the seed's list;
//   track   the persistent grid: waves take 64-particle tiles from their
//           queue (seeds grouped per XCD with queues=8), track each for at
//           most K turns, steal from other queues when theirs is drained;
// optionally captured once into a HIP graph.  Everything is enqueued on
// the caller's stream; nothing here synchronises except select_sync().
//
// run_balanced(U) is the alternative for one request on a host thread that
// may wait: every launch gives each resident wave at most one tile, taken
// from the particles furthest behind first, and the host reads the counts
// back between groups of launches.  With T tiles and W resident waves, W < T < 2W,
// the persistent form ends every chunk with T - W tiles on a nearly idle
// chip; here every launch but the last of a request fills the W waves.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "batch.h"

extern "C" void st_internal_set_error( char const* msg );

namespace st_batch
{
inline bool hip_ok( hipError_t e, const char* what )
{
    if( e == hipSuccess ) return true;
    std::string m = std::string( what ) + ": " + hipGetErrorString( e );
    st_internal_set_error( m.c_str() );
    return false;
}

inline bool fail( const char* what )
{
    st_internal_set_error( what );
    return false;
}

struct Config
{
    bool fma = false;
    int block = 256;
    int waves_per_cu = 32;
    int queues = 8;
    int64_t turns_per_launch = 0;
    bool graph = true;
    int ppl = 1;                 // particles per lane (1 or 2)
    bool ppl_auto = false;       // balanced form: per request, 2 when the particles
                                 // fill at least 70 % of the resident waves with
                                 // 128-particle tiles, else 1 (see run_balanced)
    bool mix = false;            // balanced form: mixed launches (mix_shape) where they
                                 // apply; on with ppl=auto, in either build
    bool heavy = false;          // lattices hold maps outside the lean set
                                 // (set by the owner, one particle per lane)
    bool balance = false;        // run_balanced() available (split lists)
    bool wave_prio = true;       // waves with more turns left issue first
    int staging = 0;             // lattice read through scalar loads (0) or staged
                                 // through LDS by 256-lane workgroups (1; lean
                                 // lattices, one particle per lane)

    std::string str() const
    {
        char b[ 224 ];
        std::snprintf( b, sizeof( b ), "fma=%d block=%d waves_per_cu=%d queues=%d "
                       "turns_per_launch=%lld graph=%d ppl=%s wave_prio=%d staging=%s%s%s",
                       fma ? 1 : 0, block, waves_per_cu, queues, ( long long )turns_per_launch,
                       graph ? 1 : 0, ppl_auto ? "auto" : ( ppl == 2 ? "2" : "1" ),
                       wave_prio ? 1 : 0, staging ? "lds" : "scalar",
                       heavy ? " heavy" : "", balance ? " balance" : "" );
        return b;
    }

    // one "key=value" token; unknown keys are ignored (false only on a bad value)
    bool parse_token( std::string const& k, std::string const& v )
    {
        long const iv = std::strtol( v.c_str(), nullptr, 10 );
        if( k == "fma" ) fma = iv != 0;
        else if( k == "block" )
        {
            if( iv != 64 && iv != 128 && iv != 256 ) return fail( "block must be 64, 128 or 256" );
            block = ( int )iv;
        }
        else if( k == "waves_per_cu" )
        {
            if( iv < 1 || iv > 32 ) return fail( "waves_per_cu must be 1..32" );
            waves_per_cu = ( int )iv;
        }
        else if( k == "queues" )
        {
            if( iv != 1 && iv != 8 ) return fail( "queues must be 1 or 8" );
            queues = ( int )iv;
        }
        else if( k == "turns_per_launch" ) turns_per_launch = iv > 0 ? iv : 0;
        else if( k == "graph" ) graph = iv != 0;
        else if( k == "wave_prio" ) wave_prio = iv != 0;
        else if( k == "balance" ) balance = iv != 0;
        else if( k == "ppl" )
        {
            if( v == "auto" ) { ppl_auto = true; mix = true; ppl = 1; return true; }
            if( iv != 1 && iv != 2 ) return fail( "ppl (particles per lane) must be 1, 2 or auto" );
            ppl = ( int )iv;
            ppl_auto = false;
            mix = false;
        }
        else if( k == "mix" ) mix = iv != 0;
        else if( k == "staging" )
        {
            if( v == "scalar" || v == "0" ) staging = 0;
            else if( v == "lds" || v == "1" ) staging = 1;
            else return fail( "staging must be scalar or lds" );
        }
        return true;
    }

    // the LDS-staged kernel runs lean lattices with one particle per lane
    bool lds() const { return staging != 0 && !heavy; }
    // particles per tile and tiles a workgroup tracks at once
    int64_t tile_particles() const { return lds() ? 256 : 64 * ppl; }
    int tiles_per_block() const { return lds() ? 1 : block / 64; }

    bool parse( const char* cfg )
    {
        if( cfg == nullptr ) return true;
        std::string s = cfg;
        for( char& ch : s ) if( ch == ',' || ch == ';' ) ch = ' ';
        size_t pos = 0;
        while( pos < s.size() )
        {
            size_t const end = s.find( ' ', pos );
            std::string const tok = s.substr( pos, end == std::string::npos ? std::string::npos : end - pos );
            pos = ( end == std::string::npos ) ? s.size() : end + 1;
            size_t const eq = tok.find( '=' );
            if( eq == std::string::npos ) continue;
            if( !parse_token( tok.substr( 0, eq ), tok.substr( eq + 1 ) ) ) return false;
        }
        return true;
    }
};

// Seeds -> queues without allocating (S <= ST_QUEUE_PLAN_MAX_SEEDS):
// queues = 8: greedy balance of particle counts over 8 queues (largest seed
// first), one queue per XCD; blocks are handed to the XCDs round robin, so
// block b runs on XCD b % 8 and starts on queue b % 8.  queues = 1: all
// seeds in one queue, largest first.  qs receives S seed indices, qb the
// nq + 1 queue starts.
constexpr int64_t ST_QUEUE_PLAN_MAX_SEEDS = 4096;

inline bool plan_queues( const int64_t* n_part, int64_t S, int nq, int32_t* qs, int32_t* qb )
{
    if( S > ST_QUEUE_PLAN_MAX_SEEDS || nq < 1 || nq > 8 ) return fail( "too many seeds for the queue plan" );
    int32_t order[ ST_QUEUE_PLAN_MAX_SEEDS ];
    for( int64_t s = 0; s < S; ++s ) order[ s ] = ( int32_t )s;
    std::stable_sort( order, order + S,
                      [ & ]( int32_t a, int32_t b ) { return n_part[ a ] > n_part[ b ]; } );
    int32_t queue_of[ ST_QUEUE_PLAN_MAX_SEEDS ];
    int64_t load[ 8 ] = { 0, 0, 0, 0, 0, 0, 0, 0 };
    for( int64_t r = 0; r < S; ++r )
    {
        int k = 0;
        for( int j = 1; j < nq; ++j ) if( load[ j ] < load[ k ] ) k = j;
        queue_of[ r ] = k;
        load[ k ] += n_part[ order[ r ] ];
    }
    int32_t n = 0;
    for( int k = 0; k < nq; ++k )
    {
        qb[ k ] = n;
        for( int64_t r = 0; r < S; ++r ) if( queue_of[ r ] == k ) qs[ n++ ] = order[ r ];
    }
    qb[ nq ] = n;
    return true;
}

struct BatchRunner
{
    Config cfg;
    int n_cu = 0;
    bool borrowed = false;             // arrays carved from a caller's region (init_borrowed)
    std::vector< int64_t > n_part, n_elem;
    int64_t total_particles = 0, max_particles = 0;

    st_batch_seed_t* d_seeds = nullptr;
    int32_t* d_lists = nullptr;
    int32_t* d_counts = nullptr;       // [2 * S]
    int32_t* d_qseeds = nullptr;
    int32_t* d_qbegin = nullptr;
    int32_t* d_qnext = nullptr;        // [8]
    int32_t* h_counts = nullptr;       // pinned [2 * S]
    int32_t* d_split = nullptr;        // balanced form: [2 * S] front / back counts, then gmin
    int32_t* h_split = nullptr;        // pinned [2 * S + 1]
    unsigned long long* d_progress = nullptr;   // balanced form: turns done in a launch
    hipGraphExec_t gexec = nullptr;
    std::string gkey;

    BatchRunner() = default;
    BatchRunner( const BatchRunner& ) = delete;
    BatchRunner& operator=( const BatchRunner& ) = delete;
    ~BatchRunner() { release(); }

    void release()
    {
        if( gexec ) ( void )hipGraphExecDestroy( gexec );
        gexec = nullptr; gkey.clear();
        if( !borrowed )
        {
            for( void* p : { ( void* )d_seeds, ( void* )d_lists, ( void* )d_counts, ( void* )d_qseeds,
                             ( void* )d_qbegin, ( void* )d_qnext, ( void* )d_split, ( void* )d_progress } )
                if( p ) ( void )hipFree( p );
            if( h_counts ) ( void )hipHostFree( h_counts );
            if( h_split ) ( void )hipHostFree( h_split );
        }
        borrowed = false;
        d_seeds = nullptr; d_lists = d_counts = d_qseeds = d_qbegin = d_qnext = h_counts = nullptr;
        d_split = h_split = nullptr;
        d_progress = nullptr;
        n_part.clear(); n_elem.clear(); total_particles = max_particles = 0;
    }

    int64_t n_seeds() const { return ( int64_t )n_part.size(); }

    // seeds: particle views, lattices and n_elem set; their list fields are
    // assigned here (one list region per seed)
    bool init( std::vector< st_batch_seed_t > seeds, hipStream_t stream )
    {
        release();
        int dev = 0;
        if( !hip_ok( hipGetDevice( &dev ), "hipGetDevice" ) ||
            !hip_ok( hipDeviceGetAttribute( &n_cu, hipDeviceAttributeMultiprocessorCount, dev ),
                     "hipDeviceGetAttribute" ) ) return false;
        int64_t const S = ( int64_t )seeds.size();
        for( const st_batch_seed_t& s : seeds )
        {
            if( s.P.n < 0 || s.P.n >= ( int64_t )INT32_MAX ) return fail( "particle set too large" );
            n_part.push_back( s.P.n );
            n_elem.push_back( s.n_elem );
            total_particles += s.P.n;
            max_particles = std::max( max_particles, s.P.n );
        }
        if( !hip_ok( hipMalloc( &d_lists, std::max< int64_t >( total_particles, 1 ) * sizeof( int32_t ) ),
                     "hipMalloc lists" ) ||
            !hip_ok( hipMalloc( &d_seeds, std::max< int64_t >( S, 1 ) * sizeof( st_batch_seed_t ) ),
                     "hipMalloc seeds" ) ||
            !hip_ok( hipMalloc( &d_counts, std::max< int64_t >( S, 1 ) * 2 * sizeof( int32_t ) ),
                     "hipMalloc counts" ) ||
            !hip_ok( hipHostMalloc( &h_counts, std::max< int64_t >( S, 1 ) * 2 * sizeof( int32_t ) ),
                     "hipHostMalloc counts" ) ||
            !hip_ok( hipMalloc( &d_qnext, 8 * sizeof( int32_t ) ), "hipMalloc queues" ) ) return false;
        // the balanced form's counts (a configuration may switch to it later)
        if( !hip_ok( hipMalloc( &d_split, ( 2 * std::max< int64_t >( S, 1 ) + 1 ) * sizeof( int32_t ) ),
                     "hipMalloc split counts" ) ||
            !hip_ok( hipHostMalloc( &h_split, ( 2 * std::max< int64_t >( S, 1 ) + 1 ) * sizeof( int32_t ) ),
                     "hipHostMalloc split counts" ) ||
            !hip_ok( hipMalloc( &d_progress, sizeof( unsigned long long ) ), "hipMalloc progress" ) )
            return false;
        int64_t list_off = 0;
        for( int64_t s = 0; s < S; ++s )
        {
            seeds[ s ].list = d_lists + list_off;
            list_off += n_part[ s ];
        }
        if( S > 0 && !hip_ok( hipMemcpyAsync( d_seeds, seeds.data(), S * sizeof( st_batch_seed_t ),
                                              hipMemcpyHostToDevice, stream ), "upload seeds" ) )
            return false;
        return build_queues( stream );
    }

    // No-allocation set-up (the job's NS(TrackJob_track_until_noalloc)):
    // every device array of the runner is carved out of ``d_region``
    // (region_bytes, 256-byte aligned, contents arbitrary on entry: what
    // must start at zero is cleared by the reset kernel before use) and the
    // host read-back arrays out of ``h_region`` (pinned, host_ints(S)
    // int32).  ``h_stage`` (pinned, stage_bytes(S)) stages the seed
    // descriptors and queues for their upload; it must stay untouched until
    // the stream has run the copies.  n_part / n_elem must already have
    // capacity S.  Nothing is allocated or freed.
    static constexpr size_t a256( size_t n ) { return ( n + 255u ) & ~( size_t )255u; }
    static constexpr size_t region_bytes_for( int64_t S, int64_t total )
    {
        size_t const s = ( size_t )std::max< int64_t >( S, 1 );
        return a256( s * sizeof( st_batch_seed_t ) ) + a256( 2 * s * 4 ) + a256( s * 4 ) +
               a256( 9 * 4 ) + a256( 8 * 4 ) + a256( ( 2 * s + 1 ) * 4 ) + a256( 8 ) +
               a256( ( size_t )std::max< int64_t >( total, 1 ) * 4 );
    }
    static size_t host_ints( int64_t S ) { return ( size_t )( 4 * std::max< int64_t >( S, 1 ) + 1 ); }
    static size_t stage_bytes( int64_t S )
    {
        size_t const s = ( size_t )std::max< int64_t >( S, 1 );
        return a256( s * sizeof( st_batch_seed_t ) ) + a256( ( s + 9 ) * 4 );
    }

    bool init_borrowed( const st_dev_particles_t* views, const uint64_t* lat, int64_t n_elem_per_set,
                        int64_t S, unsigned char* d_region, size_t region_bytes, int32_t* h_region,
                        unsigned char* h_stage, int device_cu, hipStream_t stream )
    {
        release();
        borrowed = true;
        n_cu = device_cu;
        if( S < 1 || S > ST_QUEUE_PLAN_MAX_SEEDS ) return fail( "no-allocation tracking: bad set count" );
        if( ( int64_t )n_part.capacity() < S || ( int64_t )n_elem.capacity() < S )
            return fail( "no-allocation tracking: runner not prepared" );
        for( int64_t s = 0; s < S; ++s )
        {
            if( views[ s ].n < 0 || views[ s ].n >= ( int64_t )INT32_MAX ) return fail( "particle set too large" );
            n_part.push_back( views[ s ].n );
            n_elem.push_back( n_elem_per_set );
            total_particles += views[ s ].n;
            max_particles = std::max( max_particles, views[ s ].n );
        }
        if( region_bytes_for( S, total_particles ) > region_bytes )
            return fail( "no-allocation tracking: particle sets exceed the scratch region" );
        unsigned char* c = d_region;
        auto carve = [ & ]( size_t n ) { unsigned char* r = c; c += a256( n ); return r; };
        d_seeds = ( st_batch_seed_t* )carve( S * sizeof( st_batch_seed_t ) );
        d_counts = ( int32_t* )carve( 2 * S * 4 );
        d_qseeds = ( int32_t* )carve( S * 4 );
        d_qbegin = ( int32_t* )carve( 9 * 4 );
        d_qnext = ( int32_t* )carve( 8 * 4 );
        d_split = ( int32_t* )carve( ( 2 * S + 1 ) * 4 );
        d_progress = ( unsigned long long* )carve( 8 );
        d_lists = ( int32_t* )carve( ( size_t )std::max< int64_t >( total_particles, 1 ) * 4 );
        h_counts = h_region;
        h_split = h_region + 2 * S;
        st_batch_seed_t* hs = ( st_batch_seed_t* )h_stage;
        int32_t* hq = ( int32_t* )( h_stage + a256( S * sizeof( st_batch_seed_t ) ) );
        int64_t list_off = 0;
        for( int64_t s = 0; s < S; ++s )
        {
            hs[ s ].P = views[ s ];
            hs[ s ].lat = lat;
            hs[ s ].n_elem = n_elem_per_set;
            hs[ s ].list = d_lists + list_off;
            list_off += n_part[ s ];
        }
        if( !plan_queues( n_part.data(), S, cfg.queues, hq, hq + S ) ) return false;
        return hip_ok( hipMemcpyAsync( d_seeds, hs, S * sizeof( st_batch_seed_t ),
                                       hipMemcpyHostToDevice, stream ), "upload seeds" ) &&
               hip_ok( hipMemcpyAsync( d_qseeds, hq, S * 4, hipMemcpyHostToDevice, stream ),
                       "upload queues" ) &&
               hip_ok( hipMemcpyAsync( d_qbegin, hq + S, ( cfg.queues + 1 ) * 4, hipMemcpyHostToDevice,
                                       stream ), "upload queues" );
    }

    // Seeds -> queues (plan_queues), uploaded to fresh queue arrays.
    bool build_queues( hipStream_t stream )
    {
        if( borrowed ) return fail( "queues of a no-allocation runner are set at its set-up" );
        if( gexec ) { ( void )hipGraphExecDestroy( gexec ); gexec = nullptr; gkey.clear(); }
        int64_t const S = n_seeds();
        int const nq = cfg.queues;
        std::vector< int32_t > qs( std::max< int64_t >( S, 1 ) ), qb( nq + 1 );
        if( !plan_queues( n_part.data(), S, nq, qs.data(), qb.data() ) ) return false;
        qs.resize( S );
        if( d_qseeds ) { ( void )hipFree( d_qseeds ); d_qseeds = nullptr; }
        if( d_qbegin ) { ( void )hipFree( d_qbegin ); d_qbegin = nullptr; }
        if( !hip_ok( hipMalloc( &d_qseeds, std::max< size_t >( qs.size(), 1 ) * sizeof( int32_t ) ),
                     "hipMalloc queue seeds" ) ||
            !hip_ok( hipMalloc( &d_qbegin, qb.size() * sizeof( int32_t ) ), "hipMalloc queue begin" ) )
            return false;
        if( !qs.empty() &&
            !hip_ok( hipMemcpyAsync( d_qseeds, qs.data(), qs.size() * sizeof( int32_t ),
                                     hipMemcpyHostToDevice, stream ), "upload queues" ) ) return false;
        return hip_ok( hipMemcpyAsync( d_qbegin, qb.data(), qb.size() * sizeof( int32_t ),
                                       hipMemcpyHostToDevice, stream ), "upload queues" ) &&
               hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" );
    }

    st_batch_sched_t sched() const
    {
        st_batch_sched_t s{};
        s.seeds = d_seeds; s.counts = d_counts; s.queue_seeds = d_qseeds;
        s.queue_begin = d_qbegin; s.queue_next = d_qnext;
        s.split = nullptr; s.gmin = nullptr; s.progress = nullptr;
        s.n_seeds = ( int32_t )n_seeds(); s.n_queues = cfg.queues;
        s.wave_prio = cfg.wave_prio ? 1 : 0;
        return s;
    }

    st_batch_sched_t sched_split() const
    {
        st_batch_sched_t s = sched();
        s.split = d_split;
        s.gmin = d_split + 2 * n_seeds();
        s.progress = d_progress;
        return s;
    }

    // turns per chunk: about 5e11 particle-element passes per launch (25
    // turns of 1M particles through 20k elements), so that a launch's
    // start-up and drain stay negligible while particles lost during a
    // chunk stop costing lanes at the next
    int64_t chunk_turns( int64_t n_active ) const
    {
        if( cfg.turns_per_launch > 0 ) return cfg.turns_per_launch;
        int64_t ne = 1;
        for( int64_t e : n_elem ) ne = std::max( ne, e );
        double const work = ( double )std::max< int64_t >( n_active, 1 ) * ( double )ne;
        double const k = 5.0e11 / work;
        return k >= 1.0e9 ? ( int64_t )1000000000 : ( k < 1.0 ? 1 : ( int64_t )k );
    }

    // tiles the persistent grid tracks at once (one per wave, or one per
    // workgroup when the lattice is staged through LDS)
    int64_t waves() const { return ( int64_t )grid() * cfg.tiles_per_block(); }

    int grid() const
    {
        int const waves_per_block = ( cfg.lds() ? 256 : cfg.block ) / 64;
        int const g = ( n_cu > 0 ? n_cu : 256 ) * cfg.waves_per_cu / waves_per_block;
        return g < cfg.queues ? cfg.queues : g;
    }

    // the longest record (words) of the lattices; the LDS-staged kernel
    // needs every record, with the next one's header, in one stage
    int64_t max_record_words = 0;
    bool staging_ok() const
    {
        return !cfg.lds() || max_record_words + 2 <= ( int64_t )ST_LAT_STAGE_WORDS ||
               fail( "a lattice record is longer than the LDS stage (staging=lds)" );
    }

    bool enqueue_reset_select( int64_t until_turn, hipStream_t stream ) const
    {
        st_batch_sched_t const s = sched();
        return hip_ok( cfg.fma ? st_launch_batch_reset_fma( &s, stream )
                               : st_launch_batch_reset_exact( &s, stream ), "launch batch reset" ) &&
               hip_ok( cfg.fma ? st_launch_batch_select_fma( &s, max_particles, until_turn, stream )
                               : st_launch_batch_select_exact( &s, max_particles, until_turn, stream ),
                       "launch batch select" );
    }

    bool enqueue_chunks( int64_t until_turn, int64_t chunks, int64_t k, int32_t* d_status,
                         hipStream_t stream ) const
    {
        st_batch_sched_t const s = sched();
        for( int64_t c = 0; c < chunks; ++c )
        {
            if( !enqueue_reset_select( until_turn, stream ) ) return false;
            hipError_t const e = cfg.fma
                ? st_launch_batch_track_fma( &s, until_turn, k, d_status, cfg.block, grid(), cfg.ppl,
                                             cfg.heavy ? 1 : 0, cfg.lds() ? 1 : 0, stream )
                : st_launch_batch_track_exact( &s, until_turn, k, d_status, cfg.block, grid(), cfg.ppl,
                                               cfg.heavy ? 1 : 0, cfg.lds() ? 1 : 0, stream );
            if( !hip_ok( e, "launch batch track" ) ) return false;
        }
        return true;
    }

    // one select pass, read back: total particles with turns to go and the
    // smallest at_turn among them (synchronises the stream)
    bool select_sync( int64_t until_turn, hipStream_t stream, int64_t& n_active, int64_t& min_turn )
    {
        int64_t const S = n_seeds();
        n_active = 0;
        min_turn = INT64_MAX;
        if( S == 0 ) { min_turn = 0; return true; }
        if( !enqueue_reset_select( until_turn, stream ) ||
            !hip_ok( hipMemcpyAsync( h_counts, d_counts, S * 2 * sizeof( int32_t ),
                                     hipMemcpyDeviceToHost, stream ), "counts D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return false;
        for( int64_t s = 0; s < S; ++s )
        {
            n_active += h_counts[ 2 * s ];
            if( h_counts[ 2 * s ] > 0 ) min_turn = std::min< int64_t >( min_turn, h_counts[ 2 * s + 1 ] );
        }
        if( n_active == 0 ) min_turn = until_turn;
        return true;
    }

    // Balanced form: turns per launch for T tiles at r turns to go.  With
    // K turns a particle needs v = ceil(r / K) visits, and the request about
    // ceil(T v / W) launches of K turns each, the last one part-filled, and
    // each ends with about a turn of drain (waves a turn apart finishing).
    // K is the one of at most chunk_turns() of W tiles that minimises
    // launches x (K + 1) over v = vmin .. vmin + 32.  With T <= W every
    // launch holds every tile: v launches of ceil(r / v) turns, at most
    // 3/2 chunk_turns() each (fewer drains; on the 20,000-element ring with
    // 1M particles and 100-turn requests, 34 turns per launch are 3 %
    // faster than 23; 50 are faster still where few particles are lost and
    // slower where more are, their lanes idling to the end of the launch).
    int64_t equal_turns( int64_t r, int64_t n_active ) const
    {
        r = std::max< int64_t >( r, 1 );
        int64_t const kmax = std::min( std::max< int64_t >( 3 * chunk_turns( n_active ) / 2, 1 ), r );
        int64_t const v = ( r + kmax - 1 ) / kmax;
        return ( r + v - 1 ) / v;
    }

    int64_t balanced_turns( int64_t tiles, int64_t r, int64_t n_active ) const
    {
        if( cfg.turns_per_launch > 0 ) return cfg.turns_per_launch;
        r = std::max< int64_t >( r, 1 );
        int64_t const W = std::max< int64_t >( waves(), 1 );
        if( tiles <= W ) return equal_turns( r, n_active );
        int64_t const kmax = std::min( chunk_turns( W * cfg.tile_particles() ), r );
        int64_t const vmin = ( r + kmax - 1 ) / kmax;
        int64_t best_k = kmax, best_cost = INT64_MAX;
        for( int64_t v = vmin; v <= std::min( vmin + 32, r ); ++v )
        {
            int64_t const k = ( r + v - 1 ) / v;
            int64_t const visits = ( r + k - 1 ) / k;
            int64_t const launches = ( tiles * visits + W - 1 ) / W;
            int64_t const cost = launches * ( k + 1 );
            if( cost < best_cost ) { best_cost = cost; best_k = k; }
        }
        return best_k;
    }

    // Balanced form: reset and the two select passes, enqueued.
    bool enqueue_select_split( int64_t until_turn, hipStream_t stream ) const
    {
        st_batch_sched_t const s = sched_split();
        return hip_ok( cfg.fma ? st_launch_batch_reset_fma( &s, stream )
                               : st_launch_batch_reset_exact( &s, stream ), "launch batch reset" ) &&
               hip_ok( cfg.fma ? st_launch_batch_select_split_fma( &s, max_particles, until_turn, stream )
                               : st_launch_batch_select_split_exact( &s, max_particles, until_turn,
                                                                     stream ),
                       "launch batch select" );
    }

    // Balanced launches with fewer tiles than resident waves get a grid of
    // about one wave per tile, which the dispatcher spreads evenly over the
    // CUs; a full grid would leave some SIMDs with 8 of the waves that have
    // a tile and others with few.  A grid smaller than the tiles only
    // defers tiles to the next launch.
    int grid_for_tiles( int64_t tiles ) const
    {
        int64_t const per_block = cfg.tiles_per_block();
        int64_t const g = ( tiles + per_block - 1 ) / per_block;
        return ( int )std::max< int64_t >( cfg.queues, std::min< int64_t >( grid(), g ) );
    }

    bool enqueue_track_split( int64_t until_turn, int64_t k, int32_t* d_status, int g,
                              hipStream_t stream ) const
    {
        st_batch_sched_t const s = sched_split();
        hipError_t const e = cfg.fma
            ? st_launch_batch_track_fma( &s, until_turn, k, d_status, cfg.block, g, cfg.ppl,
                                         cfg.heavy ? 1 : 0, cfg.lds() ? 1 : 0, stream )
            : st_launch_batch_track_exact( &s, until_turn, k, d_status, cfg.block, g, cfg.ppl,
                                           cfg.heavy ? 1 : 0, cfg.lds() ? 1 : 0, stream );
        return hip_ok( e, "launch batch track" );
    }

    // tiles of the split lists read back by select_split_sync
    int64_t split_tiles() const
    {
        int64_t const tile = cfg.tile_particles();
        int64_t t = 0;
        for( int64_t q = 0; q < 2 * n_seeds(); ++q ) t += ( h_split[ q ] + tile - 1 ) / tile;
        return t;
    }

    // Mixed launches (st_k_batch_track_mix) for one seed whose n active
    // particles sit at one turn: ``pairs`` waves of the grid take 128
    // particles each, two per lane, and the other waves share the rest
    // evenly, one per lane.  The dispatcher places one wave of each run of
    // 256 consecutive workgroups of 256 on every SIMD, so a grid of w * n_cu
    // workgroups holds w waves on every SIMD, the pair waves m of them.
    //   * More than one particle per resident lane (n > 64 W, W = 32 waves
    //     per CU): j = ceil(n / (64 x 4 x n_cu)) lanes of 64 per SIMD hold n
    //     (a pair wave counts two), and of the shapes with w + m = j the one
    //     with the most pairs and at least one single wave, w = ceil((j +
    //     1) / 2), is the fastest (1-2 % above 8 waves with the fewest
    //     pairs); used when n fills at least 90 % of the lanes (part-filled
    //     single waves cost as much as full ones).  For j = 16 every wave
    //     is a pair wave and n is spread evenly over them.
    //   * At most one per resident lane: j = ceil(n / (64 x 4 x n_cu))
    //     waves per SIMD hold n one per lane, and one wave per 64-particle
    //     tile would leave some SIMDs with j waves and the others with
    //     j - 1, idle once their waves end.  For 2 <= j <= 7 the launch
    //     spreads n evenly over j waves on every SIMD (m = 0); for j = 8,
    //     over 6 waves of which m = 2 pairs (the same lanes; up to 5 %
    //     faster than one wave per tile in the FMA build and 8 % in the
    //     exact build where the tiles leave lanes empty, equal where they
    //     fill them).
    // false when the launch does not apply (several seeds, particles at
    // several turns, a grid other than 8 workgroups of 256 per CU, at most
    // one wave per SIMD).
    bool mix_shape( int64_t n, int64_t& pairs, int& grid_wgs ) const
    {
        pairs = 0;
        grid_wgs = 0;
        if( !cfg.mix || cfg.heavy || cfg.lds() || n_seeds() != 1 ||
            h_split == nullptr || h_split[ 1 ] != 0 || cfg.block != 256 || cfg.waves_per_cu != 32 ||
            n_cu <= 0 || n <= 0 ) return false;
        int64_t const W = waves();
        if( W != 32 * ( int64_t )n_cu ) return false;
        int64_t const per = 64 * 4 * ( int64_t )n_cu;          // one wave per SIMD, one per lane
        int64_t const j = ( n + per - 1 ) / per;
        if( n > 64 * W && n <= 128 * W && j == 16 )
        {
            // more than 15 pair waves' worth per SIMD: every wave a pair
            // wave, n spread evenly over them (128-particle tiles would
            // leave whole waves idle)
            pairs = W;
            grid_wgs = grid();
            return true;
        }
        if( n > 64 * W )
        {
            int64_t const w = ( j + 2 ) / 2, m = j - w;
            if( w > 8 || 100 * n < 90 * j * per ) return false;
            pairs = m * 4 * ( int64_t )n_cu;
            grid_wgs = ( int )( w * n_cu );
            return true;
        }
        if( j < 2 ) return false;
        int64_t const m = j == 8 ? 2 : 0;
        pairs = m * 4 * ( int64_t )n_cu;
        grid_wgs = ( int )( ( j - m ) * n_cu );
        return true;
    }

    bool enqueue_track_mix( int64_t until_turn, int64_t k, int32_t* d_status, int64_t pairs,
                            int grid_wgs, hipStream_t stream ) const
    {
        st_batch_sched_t const s = sched_split();
        hipError_t const e = cfg.fma
            ? st_launch_batch_track_mix_fma( &s, until_turn, k, d_status, grid_wgs, pairs, stream )
            : st_launch_batch_track_mix_exact( &s, until_turn, k, d_status, grid_wgs, pairs, stream );
        return hip_ok( e, "launch batch track (mixed)" );
    }

    // Balanced form, one select pass read back (synchronises the stream):
    // particles with turns to go, the smallest at_turn, and the tiles of
    // the split lists.
    bool select_split_sync( int64_t until_turn, hipStream_t stream, int64_t& n_active,
                            int64_t& min_turn, int64_t& tiles )
    {
        int64_t const S = n_seeds();
        n_active = 0; tiles = 0; min_turn = until_turn;
        if( S == 0 ) return true;
        if( d_split == nullptr ) return fail( "balanced scheduling was not set up" );
        int64_t const tile = cfg.tile_particles();
        if( !enqueue_select_split( until_turn, stream ) ||
            !hip_ok( hipMemcpyAsync( h_split, d_split, ( 2 * S + 1 ) * sizeof( int32_t ),
                                     hipMemcpyDeviceToHost, stream ), "counts D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return false;
        for( int64_t k = 0; k < 2 * S; ++k )
        {
            n_active += h_split[ k ];
            tiles += ( h_split[ k ] + tile - 1 ) / tile;
        }
        if( n_active > 0 ) min_turn = h_split[ 2 * S ];
        return true;
    }

    // until_turn for every seed, balanced (see the top of this file); the
    // OR of statuses into *d_status.  The counts are read back once per
    // group of launches: a group is as many launches as the tiles need
    // without losses (every particle ceil(r / K) tile visits, W visits per
    // launch), each preceded by its select passes on the device; a group
    // that leaves particles behind (partly filled tiles) is followed by
    // another.  Every launch advances at least one tile by a turn or loses
    // its particles, so the loop ends; the bound on launches only turns a
    // fault into an error.
    bool run_balanced( int64_t until_turn, int32_t* d_status, hipStream_t stream )
    {
        if( !staging_ok() ) return false;
        int64_t n_active = 0, min_turn = 0, tiles = 0, k = 0, limit = 0, launches = 0;
        bool decide = true;                    // choose the launch form for this group
        int64_t const W = std::max< int64_t >( waves(), 1 );
        for( int64_t group = 0;; ++group )
        {
            if( !select_split_sync( until_turn, stream, n_active, min_turn, tiles ) ) return false;
            if( n_active == 0 ) return true;
            bool mixed = false;                // a mixed launch (mix_shape)
            int64_t pairs = 0;
            int mix_grid = 0;
            if( decide )
            {
                // ppl=auto: two particles per lane run the record walk once for
                // both (8 % more particle-turns per second on the 20,000-element
                // ring with 75 % or more of the resident waves busy), but hold
                // half the waves of one per lane (at 65 % they are 4 % slower);
                // they are chosen when the particles fill at least 70 % of the
                // resident waves with 128-particle tiles.  A mixed launch
                // (mix_shape) spreads the particles evenly over the waves of
                // every SIMD instead where it applies; it is decided afresh
                // before each mixed launch, as particles are lost.
                if( ( cfg.ppl_auto || cfg.mix ) && !cfg.heavy && !cfg.lds() )
                {
                    if( cfg.ppl_auto ) cfg.ppl = ( 100 * n_active >= 70 * 128 * W ) ? 2 : 1;
                    tiles = split_tiles();
                    mixed = mix_shape( n_active, pairs, mix_grid );
                }
                int64_t const r = until_turn - min_turn;
                k = mixed ? equal_turns( r, n_active ) : balanced_turns( tiles, r, n_active );
                if( group == 0 ) limit = 16 + 4 * ( ( tiles + W - 1 ) / W ) * ( ( r + k - 1 ) / k );
                decide = mixed;
            }
            if( mixed )
            {
                // one mixed launch per group: every active particle advances k
                // turns, and the next group decides again
                if( launches++ >= limit ) return fail( "balanced tracking made no progress" );
                if( !enqueue_track_mix( until_turn, k, d_status, pairs, mix_grid, stream ) ) return false;
                continue;
            }
            int64_t const v = ( until_turn - min_turn + k - 1 ) / k;     // visits per particle
            int64_t const n = std::max( v, ( tiles * v + W - 1 ) / W );
            int const g = grid_for_tiles( tiles );
            for( int64_t j = 0; j < n; ++j, ++launches )
            {
                if( launches >= limit ) return fail( "balanced tracking made no progress" );
                if( j > 0 && !enqueue_select_split( until_turn, stream ) ) return false;
                if( !enqueue_track_split( until_turn, k, d_status, g, stream ) ) return false;
            }
        }
    }

    // C chunks of K turns towards until_turn; the OR of statuses into
    // *d_status.  Captured into (and replayed from) a HIP graph with
    // cfg.graph, else launched directly.
    bool run( int64_t until_turn, int64_t chunks, int64_t k, int32_t* d_status, hipStream_t stream )
    {
        if( chunks <= 0 ) return true;
        if( !staging_ok() ) return false;
        if( !cfg.graph ) return enqueue_chunks( until_turn, chunks, k, d_status, stream );
        char kb[ 96 ];
        std::snprintf( kb, sizeof( kb ), " until=%lld chunks=%lld k=%lld status=%p",
                       ( long long )until_turn, ( long long )chunks, ( long long )k, ( void* )d_status );
        std::string const key = cfg.str() + kb;
        bool ok = true;
        if( gexec == nullptr || key != gkey )
        {
            if( gexec ) { ( void )hipGraphExecDestroy( gexec ); gexec = nullptr; gkey.clear(); }
            hipGraph_t g = nullptr;
            ok = hip_ok( hipStreamBeginCapture( stream, hipStreamCaptureModeThreadLocal ),
                         "hipStreamBeginCapture" );
            bool const enq = ok && enqueue_chunks( until_turn, chunks, k, d_status, stream );
            hipError_t const ec = ok ? hipStreamEndCapture( stream, &g ) : hipSuccess;
            ok = ok && enq && hip_ok( ec, "hipStreamEndCapture" ) &&
                 hip_ok( hipGraphInstantiate( &gexec, g, nullptr, nullptr, 0 ), "hipGraphInstantiate" );
            if( g ) ( void )hipGraphDestroy( g );
            if( ok ) gkey = key;
            else if( gexec ) { ( void )hipGraphExecDestroy( gexec ); gexec = nullptr; }
        }
        return ok && hip_ok( hipGraphLaunch( gexec, stream ), "hipGraphLaunch" );
    }
};
}  // namespace st_batch

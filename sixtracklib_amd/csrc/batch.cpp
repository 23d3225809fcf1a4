// NS(TrackBatch): many (lattice, particle set) pairs tracked together.
//
// Host side of the batched form (include/sixtracklib_amd/track_batch.h).
// Device layout:
//   * every particle buffer is uploaded as one byte image relocated to its
//     device address (as the plain job does), all images in one
//     allocation; the kernels see one st_dev_particles_t view per seed;
//   * every lattice is packed (lattice.h) into one shared allocation;
//   * per seed an active list (capacity = its particle count) and a
//     [count, smallest at_turn] pair; per queue a tile counter.
// A track_until(U) call runs C = ceil((U - t0) / K) chunks, t0 the smallest
// at_turn with turns to go, each chunk [reset, select, persistent track]
// advancing every active particle by at most K turns; the C chunks are
// captured once into a HIP graph and replayed as one launch.  The reset is
// a kernel (no memset node in the graph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "sixtracklib_amd/track_batch.h"
#include "sixtracklib_amd/particles.h"
#include "sixtracklib_amd/lattice_pack.h"
#include "lattice.h"
#include "batch.h"

extern "C" void st_internal_set_error( char const* msg );

namespace
{
bool hip_ok( hipError_t e, const char* what )
{
    if( e == hipSuccess ) return true;
    std::string m = std::string( what ) + ": " + hipGetErrorString( e );
    st_internal_set_error( m.c_str() );
    return false;
}

bool fail( const char* what )
{
    st_internal_set_error( what );
    return false;
}

size_t align256( size_t n ) { return ( n + 255u ) & ~( size_t )255u; }

struct Config
{
    bool fma = false;
    int block = 256;
    int waves_per_cu = 32;
    int queues = 8;
    int64_t turns_per_launch = 0;
    bool graph = true;

    std::string str() const
    {
        char b[ 160 ];
        std::snprintf( b, sizeof( b ), "fma=%d block=%d waves_per_cu=%d queues=%d "
                       "turns_per_launch=%lld graph=%d", fma ? 1 : 0, block, waves_per_cu,
                       queues, ( long long )turns_per_launch, graph ? 1 : 0 );
        return b;
    }

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
            std::string const k = tok.substr( 0, eq ), v = tok.substr( eq + 1 );
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
        }
        return true;
    }
};
}  // namespace

struct NS(TrackBatch)
{
    std::vector< NS(Buffer)* > lat_bufs, part_bufs;
    std::vector< size_t > img_off, img_bytes;
    std::vector< int64_t > n_part, n_elem;
    int64_t total_particles = 0, max_particles = 0;
    int64_t min_turn = 0;          // smallest at_turn among particles with turns to go
    int n_cu = 0;
    Config cfg;

    unsigned char* d_img = nullptr;
    uint64_t* d_lat = nullptr;
    int32_t* d_lists = nullptr;
    st_batch_seed_t* d_seeds = nullptr;
    int32_t* d_counts = nullptr;
    int32_t* d_qseeds = nullptr;
    int32_t* d_qbegin = nullptr;
    int32_t* d_qnext = nullptr;
    int32_t* d_status = nullptr;
    int32_t* h_status = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;

    hipGraphExec_t gexec = nullptr;
    std::string gkey;

    std::vector< st_batch_seed_t > seeds_h;

    ~NS(TrackBatch)()
    {
        if( gexec ) ( void )hipGraphExecDestroy( gexec );
        for( void* p : { ( void* )d_img, ( void* )d_lat, ( void* )d_lists, ( void* )d_seeds,
                         ( void* )d_counts, ( void* )d_qseeds, ( void* )d_qbegin,
                         ( void* )d_qnext, ( void* )d_status } )
            if( p ) ( void )hipFree( p );
        if( h_status ) ( void )hipHostFree( h_status );
        if( ev0 ) ( void )hipEventDestroy( ev0 );
        if( ev1 ) ( void )hipEventDestroy( ev1 );
        if( stream ) ( void )hipStreamDestroy( stream );
    }

    int64_t n_seeds() const { return ( int64_t )part_bufs.size(); }

    // -- setup -------------------------------------------------------------
    bool setup()
    {
        int dev = 0;
        if( !hip_ok( hipGetDevice( &dev ), "hipGetDevice" ) ||
            !hip_ok( hipDeviceGetAttribute( &n_cu, hipDeviceAttributeMultiprocessorCount, dev ),
                     "hipDeviceGetAttribute" ) ) return false;
        if( !hip_ok( hipStreamCreateWithFlags( &stream, hipStreamNonBlocking ), "hipStreamCreate" ) ||
            !hip_ok( hipEventCreate( &ev0 ), "hipEventCreate" ) ||
            !hip_ok( hipEventCreate( &ev1 ), "hipEventCreate" ) ) return false;

        int64_t const S = n_seeds();
        // particle images
        size_t total_img = 0;
        for( int64_t s = 0; s < S; ++s )
        {
            NS(Buffer)* b = part_bufs[ s ];
            if( b == nullptr || NS(Buffer_needs_remapping)( b ) ) return fail( "particle buffer invalid" );
            NS(Particles)* p = NS(Particles_buffer_get_particles)( b, 0 );
            if( p == nullptr ) return fail( "particle buffer holds no particle set" );
            if( p->num_particles >= ( int64_t )INT32_MAX ) return fail( "particle set too large" );
            img_off.push_back( total_img );
            img_bytes.push_back( NS(Buffer_get_size)( b ) );
            total_img += align256( NS(Buffer_get_size)( b ) );
            n_part.push_back( p->num_particles );
            total_particles += p->num_particles;
            max_particles = std::max( max_particles, ( int64_t )p->num_particles );
        }
        // lattices
        std::vector< std::vector< uint64_t > > packed( S );
        size_t total_lat = 0;
        std::vector< size_t > lat_off( S );
        for( int64_t s = 0; s < S; ++s )
        {
            NS(Buffer)* e = lat_bufs[ s ];
            if( e == nullptr || NS(Buffer_needs_remapping)( e ) ) return fail( "lattice buffer invalid" );
            const unsigned char* eb = NS(Buffer_get_const_data_begin)( e );
            int64_t ne = 0, nm = 0;
            int64_t const nw = NS(Lattice_pack)( eb, nullptr, 0, nullptr, &ne, nullptr, &nm );
            if( nw < 0 ) return fail( "lattice could not be packed" );
            std::vector< uint32_t > offs( ne + 1 );
            std::vector< uint32_t > mons( nm > 0 ? nm : 1 );
            packed[ s ].assign( nw + 32, 0u );
            if( NS(Lattice_pack)( eb, packed[ s ].data(), nw, offs.data(), &ne, mons.data(), &nm ) != nw )
                return fail( "lattice could not be packed" );
            for( int64_t i = 0; i < ne; ++i )
                if( !ST_OP_IS_LEAN( ST_REC_OP( packed[ s ][ offs[ i ] ] ) ) )
                    return fail( "batched tracking needs lean lattices (no monitors, beam-beam, "
                                 "space charge or RF multipoles)" );
            n_elem.push_back( ne );
            lat_off[ s ] = total_lat;
            total_lat += packed[ s ].size();
            total_lat = ( total_lat + 31u ) & ~( size_t )31u;         // 256-byte aligned streams
        }
        if( !hip_ok( hipMalloc( &d_img, std::max< size_t >( total_img, 256 ) ), "hipMalloc images" ) ||
            !hip_ok( hipMalloc( &d_lat, std::max< size_t >( total_lat, 32 ) * 8u ), "hipMalloc lattices" ) ||
            !hip_ok( hipMalloc( &d_lists, std::max< int64_t >( total_particles, 1 ) * sizeof( int32_t ) ),
                     "hipMalloc lists" ) ||
            !hip_ok( hipMalloc( &d_seeds, std::max< int64_t >( S, 1 ) * sizeof( st_batch_seed_t ) ),
                     "hipMalloc seeds" ) ||
            !hip_ok( hipMalloc( &d_counts, std::max< int64_t >( S, 1 ) * 2 * sizeof( int32_t ) ),
                     "hipMalloc counts" ) ||
            !hip_ok( hipMalloc( &d_qnext, 8 * sizeof( int32_t ) ), "hipMalloc queues" ) ||
            !hip_ok( hipMalloc( &d_status, sizeof( int32_t ) ), "hipMalloc status" ) ||
            !hip_ok( hipHostMalloc( &h_status, sizeof( int32_t ) ), "hipHostMalloc" ) ) return false;
        for( int64_t s = 0; s < S; ++s )
            if( !hip_ok( hipMemcpyAsync( d_lat + lat_off[ s ], packed[ s ].data(), packed[ s ].size() * 8u,
                                         hipMemcpyHostToDevice, stream ), "upload lattice" ) ) return false;
        // seed descriptors
        seeds_h.resize( S );
        int64_t list_off = 0;
        for( int64_t s = 0; s < S; ++s )
        {
            st_batch_seed_t& d = seeds_h[ s ];
            d.P = view( s );
            d.lat = d_lat + lat_off[ s ];
            d.n_elem = n_elem[ s ];
            d.list = d_lists + list_off;
            list_off += n_part[ s ];
        }
        if( !hip_ok( hipMemcpyAsync( d_seeds, seeds_h.data(), S * sizeof( st_batch_seed_t ),
                                     hipMemcpyHostToDevice, stream ), "upload seeds" ) ) return false;
        if( !build_queues() ) return false;
        return push();
    }

    st_dev_particles_t view( int64_t s ) const
    {
        st_dev_particles_t v{};
        NS(Buffer)* host = part_bufs[ s ];
        NS(Particles)* hp = NS(Particles_buffer_get_particles)( host, 0 );
        uint64_t const hb = NS(Buffer_get_data_begin_addr)( host );
        unsigned char* d = d_img + img_off[ s ];
        v.n = hp->num_particles;
        double* const* hr = &hp->q0;
        double** dr = &v.q0;
        for( int k = 0; k < 17; ++k )
            dr[ k ] = reinterpret_cast< double* >( d + ( ( uint64_t )( uintptr_t )hr[ k ] - hb ) );
        int64_t* const* hi = &hp->particle_id;
        int64_t** di = &v.particle_id;
        for( int k = 0; k < 4; ++k )
            di[ k ] = reinterpret_cast< int64_t* >( d + ( ( uint64_t )( uintptr_t )hi[ k ] - hb ) );
        return v;
    }

    // Seeds -> queues.  queues=8: greedy balance of particle counts over 8
    // queues (largest seed first), one queue per XCD (blocks are handed to
    // XCDs round robin, block b runs on XCD b % 8, and a block starts on
    // queue b % 8).  queues=1: all seeds in one queue, largest first.
    bool build_queues()
    {
        int64_t const S = n_seeds();
        int const nq = cfg.queues;
        std::vector< int64_t > order( S );
        std::iota( order.begin(), order.end(), 0 );
        std::stable_sort( order.begin(), order.end(),
                          [ & ]( int64_t a, int64_t b ) { return n_part[ a ] > n_part[ b ]; } );
        std::vector< std::vector< int32_t > > q( nq );
        std::vector< int64_t > load( nq, 0 );
        for( int64_t s : order )
        {
            int const k = ( int )( std::min_element( load.begin(), load.end() ) - load.begin() );
            q[ k ].push_back( ( int32_t )s );
            load[ k ] += n_part[ s ];
        }
        std::vector< int32_t > qs, qb( 1, 0 );
        for( int k = 0; k < nq; ++k )
        {
            qs.insert( qs.end(), q[ k ].begin(), q[ k ].end() );
            qb.push_back( ( int32_t )qs.size() );
        }
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
        s.n_seeds = ( int32_t )n_seeds(); s.n_queues = cfg.queues;
        return s;
    }

    // -- host <-> device -----------------------------------------------------
    bool push()
    {
        int64_t mt = INT64_MAX;
        for( int64_t s = 0; s < n_seeds(); ++s )
        {
            NS(Buffer)* b = part_bufs[ s ];
            if( NS(Buffer_get_size)( b ) != img_bytes[ s ] ) return fail( "particle buffer changed size" );
            std::vector< uint64_t > staging( ( img_bytes[ s ] + 7 ) / 8 );
            std::memcpy( staging.data(), NS(Buffer_get_const_data_begin)( b ), img_bytes[ s ] );
            unsigned char* d = d_img + img_off[ s ];
            if( NS(ManagedBuffer_remap_to)( reinterpret_cast< unsigned char* >( staging.data() ),
                                            ( NS(buffer_addr_t) )( uintptr_t )d ) != 0 )
                return fail( "relocation of particle image failed" );
            if( !hip_ok( hipMemcpyAsync( d, staging.data(), img_bytes[ s ], hipMemcpyHostToDevice,
                                         stream ), "upload particles" ) ||
                !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return false;
            NS(Particles)* p = NS(Particles_buffer_get_particles)( b, 0 );
            for( int64_t i = 0; i < p->num_particles; ++i )
                if( p->state[ i ] == 1 ) mt = std::min( mt, ( int64_t )p->at_turn[ i ] );
        }
        min_turn = ( mt == INT64_MAX ) ? 0 : mt;
        return true;
    }

    bool collect()
    {
        for( int64_t s = 0; s < n_seeds(); ++s )
        {
            NS(Buffer)* b = part_bufs[ s ];
            unsigned char* h = NS(Buffer_get_data_begin)( b );
            if( NS(Buffer_get_size)( b ) != img_bytes[ s ] ) return fail( "particle buffer changed size" );
            if( !hip_ok( hipMemcpyAsync( h, d_img + img_off[ s ], img_bytes[ s ],
                                         hipMemcpyDeviceToHost, stream ), "download particles" ) ||
                !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return false;
            if( NS(ManagedBuffer_remap_to)( h, 0u ) != 0 ) return fail( "relocation failed" );
        }
        return true;
    }

    // -- tracking ------------------------------------------------------------
    int64_t chunk_turns() const
    {
        if( cfg.turns_per_launch > 0 ) return cfg.turns_per_launch;
        int64_t ne = 1;
        for( int64_t e : n_elem ) ne = std::max( ne, e );
        double const work = ( double )std::max< int64_t >( total_particles, 1 ) * ( double )ne;
        double const k = 5.0e11 / work;
        return k >= 1.0e9 ? ( int64_t )1000000000 : ( k < 1.0 ? 1 : ( int64_t )k );
    }

    int grid() const
    {
        int const waves_per_block = cfg.block / 64;
        int const g = ( n_cu > 0 ? n_cu : 256 ) * cfg.waves_per_cu / waves_per_block;
        return g < cfg.queues ? cfg.queues : g;
    }

    bool enqueue_chunks( int64_t until_turn, int64_t chunks, int64_t k )
    {
        st_batch_sched_t const s = sched();
        for( int64_t c = 0; c < chunks; ++c )
        {
            hipError_t e1 = cfg.fma ? st_launch_batch_reset_fma( &s, stream )
                                    : st_launch_batch_reset_exact( &s, stream );
            if( !hip_ok( e1, "launch batch reset" ) ) return false;
            hipError_t e2 = cfg.fma ? st_launch_batch_select_fma( &s, max_particles, until_turn, stream )
                                    : st_launch_batch_select_exact( &s, max_particles, until_turn, stream );
            if( !hip_ok( e2, "launch batch select" ) ) return false;
            hipError_t e3 = cfg.fma
                ? st_launch_batch_track_fma( &s, until_turn, k, d_status, cfg.block, grid(), stream )
                : st_launch_batch_track_exact( &s, until_turn, k, d_status, cfg.block, grid(), stream );
            if( !hip_ok( e3, "launch batch track" ) ) return false;
        }
        return true;
    }

    int32_t track_until( int64_t until_turn, int64_t k_override = 0 )
    {
        int64_t const k = k_override > 0 ? k_override : chunk_turns();
        int64_t const chunks = until_turn > min_turn ? ( until_turn - min_turn + k - 1 ) / k : 0;
        *h_status = 0;
        if( !hip_ok( hipMemcpyAsync( d_status, h_status, sizeof( int32_t ), hipMemcpyHostToDevice,
                                     stream ), "status reset" ) ||
            !hip_ok( hipEventRecord( ev0, stream ), "hipEventRecord" ) ) return -1;
        bool ok = true;
        if( chunks > 0 )
        {
            if( cfg.graph )
            {
                std::string const key = cfg.str() + " until=" + std::to_string( until_turn ) +
                                        " chunks=" + std::to_string( chunks ) + " k=" + std::to_string( k );
                if( gexec == nullptr || key != gkey )
                {
                    if( gexec ) { ( void )hipGraphExecDestroy( gexec ); gexec = nullptr; }
                    hipGraph_t g = nullptr;
                    ok = hip_ok( hipStreamBeginCapture( stream, hipStreamCaptureModeThreadLocal ),
                                 "hipStreamBeginCapture" );
                    bool const enq = ok && enqueue_chunks( until_turn, chunks, k );
                    hipError_t const ec = hipStreamEndCapture( stream, &g );
                    ok = ok && enq && hip_ok( ec, "hipStreamEndCapture" ) &&
                         hip_ok( hipGraphInstantiate( &gexec, g, nullptr, nullptr, 0 ),
                                 "hipGraphInstantiate" );
                    if( g ) ( void )hipGraphDestroy( g );
                    if( ok ) gkey = key;
                    else if( gexec ) { ( void )hipGraphExecDestroy( gexec ); gexec = nullptr; }
                }
                ok = ok && hip_ok( hipGraphLaunch( gexec, stream ), "hipGraphLaunch" );
            }
            else
            {
                ok = enqueue_chunks( until_turn, chunks, k );
            }
        }
        if( !ok ) return -1;
        if( !hip_ok( hipEventRecord( ev1, stream ), "hipEventRecord" ) ||
            !hip_ok( hipMemcpyAsync( h_status, d_status, sizeof( int32_t ), hipMemcpyDeviceToHost,
                                     stream ), "status D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return -1;
        float ms = 0.0f;
        last_ms = ( hipEventElapsedTime( &ms, ev0, ev1 ) == hipSuccess ) ? ms : -1.0f;
        if( chunks > 0 ) min_turn = until_turn;
        return *h_status;
    }

    // Candidate (block, waves_per_cu, queues) timed on trial_turns turns of
    // the batch itself; the particle images are saved first and restored
    // after every trial, so the batch state is unchanged.
    int autotune( int64_t trial_turns )
    {
        if( trial_turns <= 0 ) trial_turns = 2;
        size_t total_img = 0;
        for( int64_t s = 0; s < n_seeds(); ++s ) total_img = img_off[ s ] + align256( img_bytes[ s ] );
        unsigned char* backup = nullptr;
        if( !hip_ok( hipMalloc( &backup, std::max< size_t >( total_img, 256 ) ), "hipMalloc backup" ) )
            return -1;
        bool ok = hip_ok( hipMemcpyAsync( backup, d_img, total_img, hipMemcpyDeviceToDevice, stream ),
                          "backup" ) && hip_ok( hipStreamSynchronize( stream ), "sync" );
        Config const base = cfg;
        int64_t const t0 = min_turn;
        Config best = base;
        float best_ms = -1.0f;
        int const blocks[] = { 64, 128, 256 };
        int const waves[] = { 16, 32 };
        int const queues[] = { 1, 8 };
        for( int b : blocks )
            for( int w : waves )
                for( int q : queues )
                {
                    if( !ok ) break;
                    cfg = base;
                    cfg.block = b; cfg.waves_per_cu = w; cfg.queues = q; cfg.graph = false;
                    if( !build_queues() ) { ok = false; break; }
                    min_turn = t0;
                    int32_t const st = track_until( t0 + trial_turns, trial_turns );
                    ok = st >= -1 && last_ms >= 0.0f;
                    if( ok && ( best_ms < 0.0f || last_ms < best_ms ) ) { best_ms = last_ms; best = cfg; }
                    ok = ok && hip_ok( hipMemcpyAsync( d_img, backup, total_img, hipMemcpyDeviceToDevice,
                                                       stream ), "restore" ) &&
                         hip_ok( hipStreamSynchronize( stream ), "sync" );
                }
        ( void )hipFree( backup );
        cfg = best;
        cfg.graph = base.graph;
        min_turn = t0;
        if( gexec ) { ( void )hipGraphExecDestroy( gexec ); gexec = nullptr; gkey.clear(); }
        if( !build_queues() ) return -1;
        return ok ? 0 : -1;
    }
};

extern "C" {

NS(TrackBatch)* NS(TrackBatch_new)( int64_t n_seeds, NS(Buffer)* const* lattices,
    NS(Buffer)* const* particles, char const* config_str )
{
    if( n_seeds <= 0 || lattices == nullptr || particles == nullptr )
    {
        st_internal_set_error( "TrackBatch_new: no seeds" );
        return nullptr;
    }
    NS(TrackBatch)* b = new NS(TrackBatch)();
    b->lat_bufs.assign( lattices, lattices + n_seeds );
    b->part_bufs.assign( particles, particles + n_seeds );
    if( !b->cfg.parse( config_str ) || !b->setup() ) { delete b; return nullptr; }
    return b;
}

void NS(TrackBatch_delete)( NS(TrackBatch)* batch ) { delete batch; }

NS(track_status_t) NS(TrackBatch_track_until)( NS(TrackBatch)* batch, int64_t until_turn )
{
    return batch ? batch->track_until( until_turn ) : -1;
}

int NS(TrackBatch_collect)( NS(TrackBatch)* batch ) { return ( batch && batch->collect() ) ? 0 : -1; }
int NS(TrackBatch_push)( NS(TrackBatch)* batch ) { return ( batch && batch->push() ) ? 0 : -1; }

int NS(TrackBatch_set_config)( NS(TrackBatch)* batch, char const* config_str )
{
    if( batch == nullptr ) return -1;
    Config c = batch->cfg;
    if( !c.parse( config_str ) ) return -1;
    bool const requeue = c.queues != batch->cfg.queues;
    batch->cfg = c;
    if( batch->gexec ) { ( void )hipGraphExecDestroy( batch->gexec ); batch->gexec = nullptr; batch->gkey.clear(); }
    return ( !requeue || batch->build_queues() ) ? 0 : -1;
}

int64_t NS(TrackBatch_get_config)( const NS(TrackBatch)* batch, char* out, int64_t out_len )
{
    if( batch == nullptr ) return -1;
    std::string const s = batch->cfg.str();
    if( out != nullptr && out_len > 0 )
    {
        size_t const n = std::min< size_t >( s.size(), ( size_t )( out_len - 1 ) );
        std::memcpy( out, s.data(), n );
        out[ n ] = '\0';
    }
    return ( int64_t )s.size();
}

int NS(TrackBatch_autotune)( NS(TrackBatch)* batch, int64_t trial_turns )
{
    return batch ? batch->autotune( trial_turns ) : -1;
}

int64_t NS(TrackBatch_num_seeds)( const NS(TrackBatch)* batch ) { return batch ? batch->n_seeds() : -1; }
int64_t NS(TrackBatch_total_particles)( const NS(TrackBatch)* batch )
{
    return batch ? batch->total_particles : -1;
}

double NS(TrackBatch_last_kernel_time_ms)( const NS(TrackBatch)* batch )
{
    return batch ? ( double )batch->last_ms : -1.0;
}

}  // extern "C"

// NS(TrackBatch): many (lattice, particle set) pairs tracked together.
//
// Host side of the batched form (include/sixtracklib_amd/track_batch.h).
// Device layout:
//   * every particle buffer is uploaded as one byte image relocated to its
//     device address (as the plain job does), all images in one
//     allocation; the kernels see one st_dev_particles_t view per seed;
//   * every lattice is packed (lattice.h) into one shared allocation;
//   * lists, counts and queues belong to the BatchRunner (batch_runner.h).
// A track_until(U) call runs C = ceil((U - t0) / K) chunks, t0 the smallest
// at_turn with turns to go, captured once into a HIP graph and replayed as
// one launch.  t0 is known on the host: it is read from the host buffers
// at push, and after track_until(U) every particle with turns to go is at
// U.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "sixtracklib_amd/track_batch.h"
#include "sixtracklib_amd/particles.h"
#include "sixtracklib_amd/lattice_pack.h"
#include "lattice.h"
#include "batch.h"
#include "batch_runner.h"

using st_batch::hip_ok;
using st_batch::fail;
using st_batch::Config;

namespace
{
size_t align256( size_t n ) { return ( n + 255u ) & ~( size_t )255u; }
}  // namespace

struct NS(TrackBatch)
{
    std::vector< NS(Buffer)* > lat_bufs, part_bufs;
    std::vector< size_t > img_off, img_bytes;
    size_t total_img = 0;
    int64_t min_turn = 0;          // smallest at_turn among particles with turns to go
    st_batch::BatchRunner run;

    unsigned char* d_img = nullptr;
    uint64_t* d_lat = nullptr;
    int32_t* d_status = nullptr;
    int32_t* h_status = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;

    ~NS(TrackBatch)()
    {
        run.release();
        for( void* p : { ( void* )d_img, ( void* )d_lat, ( void* )d_status } )
            if( p ) ( void )hipFree( p );
        if( h_status ) ( void )hipHostFree( h_status );
        if( ev0 ) ( void )hipEventDestroy( ev0 );
        if( ev1 ) ( void )hipEventDestroy( ev1 );
        if( stream ) ( void )hipStreamDestroy( stream );
    }

    int64_t n_seeds() const { return ( int64_t )part_bufs.size(); }

    bool setup( Config const& cfg )
    {
        run.cfg = cfg;
        if( !hip_ok( hipStreamCreateWithFlags( &stream, hipStreamNonBlocking ), "hipStreamCreate" ) ||
            !hip_ok( hipEventCreate( &ev0 ), "hipEventCreate" ) ||
            !hip_ok( hipEventCreate( &ev1 ), "hipEventCreate" ) ) return false;
        int64_t const S = n_seeds();
        for( int64_t s = 0; s < S; ++s )
        {
            NS(Buffer)* b = part_bufs[ s ];
            if( b == nullptr || NS(Buffer_needs_remapping)( b ) ) return fail( "particle buffer invalid" );
            if( NS(Particles_buffer_get_particles)( b, 0 ) == nullptr )
                return fail( "particle buffer holds no particle set" );
            img_off.push_back( total_img );
            img_bytes.push_back( NS(Buffer_get_size)( b ) );
            total_img += align256( NS(Buffer_get_size)( b ) );
        }
        std::vector< std::vector< uint64_t > > packed( S );
        std::vector< int64_t > n_elem( S );
        std::vector< size_t > lat_off( S );
        size_t total_lat = 0;
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
            packed[ s ].assign( nw + ST_LAT_PAD_WORDS, 0u );   // zero padding read ahead by the kernels
            if( NS(Lattice_pack)( eb, packed[ s ].data(), nw, offs.data(), &ne, mons.data(), &nm ) != nw )
                return fail( "lattice could not be packed" );
            for( int64_t i = 0; i < ne; ++i )
                if( !ST_OP_IS_LEAN( ST_REC_OP( packed[ s ][ offs[ i ] ] ) ) )
                    return fail( "batched tracking needs lean lattices (no monitors, beam-beam, "
                                 "space charge or RF multipoles)" );
            for( int64_t i = 0; i < ne; ++i )
                run.max_record_words = std::max< int64_t >( run.max_record_words, offs[ i + 1 ] - offs[ i ] );
            n_elem[ s ] = ne;
            lat_off[ s ] = total_lat;
            total_lat += packed[ s ].size();
            total_lat = ( total_lat + 31u ) & ~( size_t )31u;         // 256-byte aligned streams
        }
        if( !hip_ok( hipMalloc( &d_img, std::max< size_t >( total_img, 256 ) ), "hipMalloc images" ) ||
            !hip_ok( hipMalloc( &d_lat, std::max< size_t >( total_lat, 32 ) * 8u ), "hipMalloc lattices" ) ||
            !hip_ok( hipMalloc( &d_status, sizeof( int32_t ) ), "hipMalloc status" ) ||
            !hip_ok( hipHostMalloc( &h_status, sizeof( int32_t ) ), "hipHostMalloc" ) ) return false;
        for( int64_t s = 0; s < S; ++s )
            if( !hip_ok( hipMemcpyAsync( d_lat + lat_off[ s ], packed[ s ].data(), packed[ s ].size() * 8u,
                                         hipMemcpyHostToDevice, stream ), "upload lattice" ) ) return false;
        std::vector< st_batch_seed_t > seeds( S );
        for( int64_t s = 0; s < S; ++s )
        {
            seeds[ s ].P = view( s );
            seeds[ s ].lat = d_lat + lat_off[ s ];
            seeds[ s ].n_elem = n_elem[ s ];
            seeds[ s ].list = nullptr;
        }
        return run.init( seeds, stream ) && push();     // both synchronise: packed outlives the copies
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
    int32_t track_until( int64_t until_turn, int64_t k_override = 0 )
    {
        last_ms = -1.0f;
        int64_t const k = k_override > 0 ? k_override : run.chunk_turns( run.total_particles );
        int64_t const chunks = until_turn > min_turn ? ( until_turn - min_turn + k - 1 ) / k : 0;
        *h_status = 0;
        if( !hip_ok( hipMemcpyAsync( d_status, h_status, sizeof( int32_t ), hipMemcpyHostToDevice,
                                     stream ), "status reset" ) ||
            !hip_ok( hipEventRecord( ev0, stream ), "hipEventRecord" ) ||
            !run.run( until_turn, chunks, k, d_status, stream ) ||
            !hip_ok( hipEventRecord( ev1, stream ), "hipEventRecord" ) ||
            !hip_ok( hipMemcpyAsync( h_status, d_status, sizeof( int32_t ), hipMemcpyDeviceToHost,
                                     stream ), "status D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return -1;
        float ms = 0.0f;
        last_ms = ( hipEventElapsedTime( &ms, ev0, ev1 ) == hipSuccess ) ? ms : -1.0f;
        if( chunks > 0 ) min_turn = until_turn;
        return *h_status;
    }

    // The balanced form (batch_runner.h, config token balance=1): launches
    // of one tile per resident wave, the host reading the counts back
    // between launch groups (no graph).
    int32_t track_until_balanced( int64_t until_turn )
    {
        last_ms = -1.0f;
        *h_status = 0;
        if( !hip_ok( hipMemcpyAsync( d_status, h_status, sizeof( int32_t ), hipMemcpyHostToDevice,
                                     stream ), "status reset" ) ||
            !hip_ok( hipEventRecord( ev0, stream ), "hipEventRecord" ) ||
            !run.run_balanced( until_turn, d_status, stream ) ||
            !hip_ok( hipEventRecord( ev1, stream ), "hipEventRecord" ) ||
            !hip_ok( hipMemcpyAsync( h_status, d_status, sizeof( int32_t ), hipMemcpyDeviceToHost,
                                     stream ), "status D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return -1;
        float ms = 0.0f;
        last_ms = ( hipEventElapsedTime( &ms, ev0, ev1 ) == hipSuccess ) ? ms : -1.0f;
        if( until_turn > min_turn ) min_turn = until_turn;
        return *h_status;
    }

    // Candidate configurations timed on trial_turns turns of the batch
    // itself, the three fastest a second time.  The trials start after
    // skip_turns turns tracked with the current configuration (default 10
    // trials' worth), so that they see the batch after its early losses,
    // not the first turns with every particle alive.  The particle images
    // are saved first and restored at the end (and the advanced state before
    // every trial), so the batch state is unchanged.  Candidates:
    //   one particle per lane, scalar-load lattice: block 64/128/256, 4 or 8
    //     resident waves per SIMD, one queue or one per XCD;
    //   two particles per lane (128-particle tiles, 5 waves per SIMD): block
    //     256, both queue layouts;
    //   lattice staged through LDS (256-lane workgroups): 4 or 8 waves per
    //     SIMD, both queue layouts.
    int autotune( int64_t trial_turns, int64_t skip_turns )
    {
        if( trial_turns <= 0 ) trial_turns = 2;
        if( skip_turns < 0 ) skip_turns = 10 * trial_turns;
        size_t const bytes = std::max< size_t >( total_img, 256 );
        unsigned char* backup = nullptr;     // the state on entry
        unsigned char* window = nullptr;     // the state the trials start from
        if( !hip_ok( hipMalloc( &backup, bytes ), "hipMalloc backup" ) ) return -1;
        if( !hip_ok( hipMalloc( &window, bytes ), "hipMalloc backup" ) )
        { ( void )hipFree( backup ); return -1; }
        Config const base = run.cfg;
        int64_t const t_entry = min_turn;
        bool ok = hip_ok( hipMemcpyAsync( backup, d_img, total_img, hipMemcpyDeviceToDevice, stream ),
                          "backup" ) && hip_ok( hipStreamSynchronize( stream ), "sync" );
        if( ok && skip_turns > 0 )
            ok = ( run.cfg.balance ? track_until_balanced( t_entry + skip_turns )
                                   : track_until( t_entry + skip_turns ) ) >= -1 && last_ms >= 0.0f;
        int64_t const t0 = ok ? std::max( min_turn, t_entry ) : t_entry;
        ok = ok && hip_ok( hipMemcpyAsync( window, d_img, total_img, hipMemcpyDeviceToDevice, stream ),
                           "backup" ) && hip_ok( hipStreamSynchronize( stream ), "sync" );
        auto trial = [ & ]( Config const& c, float& ms ) -> bool
        {
            run.cfg = c;
            if( !run.build_queues( stream ) ) return false;
            min_turn = t0;
            int32_t const st = run.cfg.balance ? track_until_balanced( t0 + trial_turns )
                                               : track_until( t0 + trial_turns, trial_turns );
            ms = last_ms;
            // the images go back to the window's start whatever the trial gave
            bool const restored =
                hip_ok( hipMemcpyAsync( d_img, window, total_img, hipMemcpyDeviceToDevice, stream ),
                        "restore" ) &&
                hip_ok( hipStreamSynchronize( stream ), "sync" );
            return restored && st >= -1 && ms >= 0.0f;
        };
        std::vector< Config > cand;
        for( int b : { 64, 128, 256 } )
            for( int w : { 16, 32 } )
                for( int q : { 1, 8 } )
                {
                    Config c = base;
                    c.ppl = 1; c.ppl_auto = false; c.staging = 0; c.block = b; c.waves_per_cu = w;
                    c.queues = q;
                    cand.push_back( c );
                }
        for( int w : { 16, 32 } )
            for( int q : { 1, 8 } )
            {
                Config c = base;
                c.ppl = 2; c.ppl_auto = false; c.staging = 0; c.block = 256; c.waves_per_cu = w;
                c.queues = q;
                cand.push_back( c );
            }
        if( run.max_record_words + 2 <= ( int64_t )ST_LAT_STAGE_WORDS )
            for( int w : { 16, 32 } )
                for( int q : { 1, 8 } )
                {
                    Config c = base;
                    c.ppl = 1; c.ppl_auto = false; c.staging = 1; c.block = 256; c.waves_per_cu = w;
                    c.queues = q;
                    cand.push_back( c );
                }
        std::vector< std::pair< float, Config > > timed;
        for( Config c : cand )
        {
            if( !ok ) break;
            c.graph = false;
            float ms = -1.0f;
            ok = trial( c, ms );
            if( ok ) timed.emplace_back( ms, c );
        }
        std::stable_sort( timed.begin(), timed.end(),
                          []( auto const& x, auto const& y ) { return x.first < y.first; } );
        Config best = base;
        float best_ms = -1.0f;
        for( size_t r = 0; ok && r < std::min< size_t >( 3, timed.size() ); ++r )
        {
            float ms = -1.0f;
            ok = trial( timed[ r ].second, ms );
            float const m = std::min( ms, timed[ r ].first );
            if( ok && ( best_ms < 0.0f || m < best_ms ) ) { best_ms = m; best = timed[ r ].second; }
        }
        // back to the state on entry
        bool const restored =
            hip_ok( hipMemcpyAsync( d_img, backup, total_img, hipMemcpyDeviceToDevice, stream ),
                    "restore" ) &&
            hip_ok( hipStreamSynchronize( stream ), "sync" );
        ( void )hipFree( backup );
        ( void )hipFree( window );
        run.cfg = ok ? best : base;
        run.cfg.graph = base.graph;
        min_turn = t_entry;
        if( !run.build_queues( stream ) ) return -1;
        return ( ok && restored ) ? 0 : -1;
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
    Config cfg;
    cfg.balance = true;                  // default: launches of one tile per resident wave
    if( !cfg.parse( config_str ) ) return nullptr;
    NS(TrackBatch)* b = new NS(TrackBatch)();
    b->lat_bufs.assign( lattices, lattices + n_seeds );
    b->part_bufs.assign( particles, particles + n_seeds );
    if( !b->setup( cfg ) ) { delete b; return nullptr; }
    return b;
}

void NS(TrackBatch_delete)( NS(TrackBatch)* batch ) { delete batch; }

NS(track_status_t) NS(TrackBatch_track_until)( NS(TrackBatch)* batch, int64_t until_turn )
{
    if( batch == nullptr ) return -1;
    return batch->run.cfg.balance ? batch->track_until_balanced( until_turn )
                                  : batch->track_until( until_turn );
}

int NS(TrackBatch_collect)( NS(TrackBatch)* batch ) { return ( batch && batch->collect() ) ? 0 : -1; }
int NS(TrackBatch_push)( NS(TrackBatch)* batch ) { return ( batch && batch->push() ) ? 0 : -1; }

int NS(TrackBatch_set_config)( NS(TrackBatch)* batch, char const* config_str )
{
    if( batch == nullptr ) return -1;
    Config c = batch->run.cfg;
    if( !c.parse( config_str ) ) return -1;
    batch->run.cfg = c;
    return batch->run.build_queues( batch->stream ) ? 0 : -1;
}

int64_t NS(TrackBatch_get_config)( const NS(TrackBatch)* batch, char* out, int64_t out_len )
{
    if( batch == nullptr ) return -1;
    std::string const s = batch->run.cfg.str();
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
    return batch ? batch->autotune( trial_turns, -1 ) : -1;
}

int NS(TrackBatch_autotune_window)( NS(TrackBatch)* batch, int64_t skip_turns, int64_t trial_turns )
{
    return batch ? batch->autotune( trial_turns, skip_turns < 0 ? 0 : skip_turns ) : -1;
}

int64_t NS(TrackBatch_num_seeds)( const NS(TrackBatch)* batch ) { return batch ? batch->n_seeds() : -1; }
int64_t NS(TrackBatch_total_particles)( const NS(TrackBatch)* batch )
{
    return batch ? batch->run.total_particles : -1;
}

double NS(TrackBatch_last_kernel_time_ms)( const NS(TrackBatch)* batch )
{
    return batch ? ( double )batch->last_ms : -1.0;
}

}  // extern "C"

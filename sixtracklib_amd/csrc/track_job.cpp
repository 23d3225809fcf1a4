// NS(TrackJob) on HIP: host side of the drop-in boundary.
//
// Reference behaviour followed:
//   construction / reset      common/internal/track_job_base.cpp:2107-2437
//   output preparation        common/output/output_buffer.c:44-363,
//                             be_monitor/output_buffer.h:96-140,
//                             output/elem_by_elem_config.h:1050-1110
//   track / collect / push    common/internal/track_job.cpp:417-446,
//                             cuda/track/track_job.cpp:2321-2390
// Design (this project's): the particle and output buffers live on the
// device as byte images of the host CObjects buffers relocated to the
// device base address, so collect is one copy + one host relocation; the
// lattice is flattened once into the packed stream of lattice.h; every
// launch runs on a job-owned stream; track_* return after the launch has
// finished and the OR'ed status has been read back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sixtracklib_amd/track_job.h"
#include "sixtracklib_amd/beam_elements.h"
#include "sixtracklib_amd/lattice_pack.h"
#include "lattice.h"
#include "track_launch.h"
#include "batch.h"
#include "batch_runner.h"

// Scratch memory of NS(TrackJob_track_until_noalloc): the scheduler's
// per-call device arrays (status word, seed descriptors, counts, queues,
// active-particle lists) for at most ST_NOALLOC_MAX_SETS particle sets of
// together at most ST_NOALLOC_MAX_PARTICLES particles.  Every GPU the
// library runs on holds one copy until the process ends.  In host code the
// name ``scratch`` is only a stand-in for the symbol: the device address
// of each GPU's copy comes from hipGetSymbolAddress (scratch_address).
constexpr int64_t ST_NOALLOC_MAX_SETS = 64;
constexpr int64_t ST_NOALLOC_MAX_PARTICLES = int64_t( 1 ) << 21;
constexpr size_t ST_NOALLOC_STATUS_BYTES = 256;
constexpr size_t ST_NOALLOC_SCRATCH_BYTES = ( size_t( 8 ) << 20 ) + ( size_t( 64 ) << 10 );

static_assert( ST_NOALLOC_STATUS_BYTES + st_batch::BatchRunner::region_bytes_for(
                   ST_NOALLOC_MAX_SETS, ST_NOALLOC_MAX_PARTICLES ) <= ST_NOALLOC_SCRATCH_BYTES,
               "scratch memory holds the largest no-allocation call" );

static __device__ __align__( 256 ) unsigned char scratch[ ST_NOALLOC_SCRATCH_BYTES ];

namespace
{
thread_local std::string g_err;

constexpr int ST_NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[ ST_NOALLOC_MAX_DEVICES ] = {};

// Device address of ``scratch`` on ``device``: looked up once per device,
// with that device current, at the set-up of the first job on it.  nullptr
// (and the error message set) when the lookup fails.
void* scratch_address( int device, bool lookup )
{
    if( device < 0 || device >= ST_NOALLOC_MAX_DEVICES ) return nullptr;
    std::lock_guard< std::mutex > lock( g_scratch_mutex );
    if( g_scratch_address[ device ] != nullptr || !lookup ) return g_scratch_address[ device ];
    int current = -1;
    if( hipGetDevice( &current ) != hipSuccess ) return nullptr;
    if( current != device && hipSetDevice( device ) != hipSuccess ) return nullptr;
    void* ptr = nullptr;
    hipError_t const e = hipGetSymbolAddress( ( void** )&ptr, HIP_SYMBOL( scratch ) );
    if( current != device ) ( void )hipSetDevice( current );
    if( e != hipSuccess || ptr == nullptr )
    {
        g_err = std::string( "hipGetSymbolAddress(scratch): " ) + hipGetErrorString( e );
        return nullptr;
    }
    g_scratch_address[ device ] = ptr;
    return ptr;
}

bool hip_ok( hipError_t e, const char* what )
{
    if( e == hipSuccess ) return true;
    g_err = std::string( what ) + ": " + hipGetErrorString( e );
    return false;
}

bool is_element_type( uint64_t t )
{
    return t != ST_OBJECT_TYPE_NONE && t != ST_OBJECT_TYPE_PARTICLE &&
           t != ST_OBJECT_TYPE_INVALID && t != ST_OBJECT_TYPE_LINE &&
           t != ST_OBJECT_TYPE_ELEM_BY_ELEM_CONF;
}

struct Attr
{
    int64_t min_pid = LLONG_MAX, max_pid = LLONG_MIN;
    int64_t min_elem = LLONG_MAX, max_elem = LLONG_MIN;
    int64_t min_turn = LLONG_MAX, max_turn = LLONG_MIN;
};

// output/output_buffer.h:461-560 (particle extrema folded with the
// element-id range of the lattice, elem_by_elem_output_buffer.h:163-235)
bool find_attributes( NS(Buffer)* pbuf, const std::vector< uint64_t >& sets,
                      NS(Buffer)* ebuf, Attr& a, int64_t* n_ebe_objs )
{
    for( uint64_t k : sets )
    {
        NS(Particles) const* p = NS(Particles_buffer_get_const_particles)( pbuf, k );
        if( p == nullptr ) return false;
        NS(Particles_find_min_max_attributes)( p, &a.min_pid, &a.max_pid,
            &a.min_elem, &a.max_elem, &a.min_turn, &a.max_turn );
    }
    if( a.min_pid == LLONG_MAX ) { a.min_pid = a.max_pid = 0; a.min_elem = a.max_elem = 0;
                                   a.min_turn = a.max_turn = 0; }
    if( a.min_pid < 0 ) a.min_pid = 0;
    int64_t const n = ( int64_t )NS(Buffer_get_num_of_objects)( ebuf );
    int64_t lo = 0, count = 0;
    int64_t i = 0;
    for( ; i < n; ++i )
    {
        if( is_element_type( NS(Buffer_get_const_object)( ebuf, i )->type_id ) ) { ++count; ++i; break; }
        ++lo;
    }
    int64_t hi = lo;
    for( ; i < n; ++i )
        if( is_element_type( NS(Buffer_get_const_object)( ebuf, i )->type_id ) ) { ++count; ++hi; }
    if( count == 0 ) lo = hi = 0;
    a.min_elem = std::min( a.min_elem, lo );
    a.max_elem = std::max( a.max_elem, hi );
    if( n_ebe_objs ) *n_ebe_objs = count;
    return true;
}

// integer value of ``key=`` in a configuration string (tokens separated by
// blanks, commas or semicolons), ``def`` when absent
long config_int( const char* cfg, const char* key, long def )
{
    if( cfg == nullptr ) return def;
    std::string s = cfg;
    for( char& ch : s ) if( ch == ',' || ch == ';' ) ch = ' ';
    std::string const k = std::string( key ) + "=";
    size_t pos = 0;
    long v = def;
    while( ( pos = s.find( k, pos ) ) != std::string::npos )
    {
        if( pos == 0 || s[ pos - 1 ] == ' ' )
            v = std::strtol( s.c_str() + pos + k.size(), nullptr, 10 );
        pos += k.size();
    }
    return v;
}

struct OutputPlan
{
    bool has_ebe = false;
    int64_t ebe_size = 0, max_ebe_turn = 0;
    int64_t ebe_cap = 0;                       // store capacity in records
    std::vector< int64_t > monitor_elems;      // element index of every monitor
    std::vector< int64_t > monitor_sizes;
};

// output/output_buffer.c:288-363.  ``ebe_passes`` > 0 sizes the
// element-by-element store for that many element passes of every particle
// when that is fewer than the reference's (elements x turns).
void plan_output( NS(Buffer)* ebuf, const Attr& a, uint64_t until_turn_ebe, int64_t ebe_passes,
                  OutputPlan& o )
{
    int64_t const np = a.max_pid - a.min_pid + 1;
    if( a.min_turn >= 0 && ( int64_t )until_turn_ebe > a.min_turn )
    {
        o.has_ebe = true;
        o.max_ebe_turn = ( int64_t )until_turn_ebe - 1;
        int64_t const passes = ( a.max_elem - a.min_elem + 1 ) * ( o.max_ebe_turn - a.min_turn + 1 );
        o.ebe_size = np * ( ( ebe_passes > 0 && ebe_passes < passes ) ? ebe_passes : passes );
        o.ebe_cap = o.ebe_size;
    }
    int64_t const n = ( int64_t )NS(Buffer_get_num_of_objects)( ebuf );
    for( int64_t i = 0; i < n; ++i )
    {
        NS(Object) const* obj = NS(Buffer_get_const_object)( ebuf, i );
        if( obj->type_id != ST_OBJECT_TYPE_BEAM_MONITOR ) continue;
        NS(BeamMonitor) const* m = ( NS(BeamMonitor) const* )( uintptr_t )obj->begin_addr;
        if( m->num_stores <= 0 ) continue;
        o.monitor_elems.push_back( i );
        o.monitor_sizes.push_back( m->num_stores * np );
    }
}
}  // namespace

struct NS(TrackJobBase)
{
    std::string arch = "hip", config, device_id;
    NS(Buffer)* pbuf = nullptr;
    NS(Buffer)* ebuf = nullptr;
    NS(Buffer)* obuf = nullptr;
    bool owns_obuf = false;
    std::vector< uint64_t > sets;
    std::vector< int64_t > set_sizes;
    Attr attr;
    OutputPlan plan;
    uint64_t ebe_offset = 0, monitor_offset = 0, until_turn_ebe = 0;
    st_ebe_config_t ebe_cfg{};
    bool ebe_rolling = true, fma = false;
    int64_t ebe_passes = 0;              // 0: the full (elements x turns) store
    int block = 256, device = 0;
    uint16_t collect_flags = ST_TRACK_JOB_COLLECT_DEFAULT_FLAGS;
    int32_t last_status = 0;
    bool debug_mode = false;             // new-style API's debug mode (changes nothing)

    hipStream_t stream = nullptr;
    unsigned char* d_pbuf = nullptr; size_t pbuf_bytes = 0;
    unsigned char* d_obuf = nullptr; size_t obuf_bytes = 0;
    uint64_t* d_lat = nullptr; int64_t lat_words = 0, n_elem = 0;
    std::vector< uint32_t > rec_offsets, monitor_fields;
    std::vector< uint64_t > lat_host;
    bool lat_full = false;
    int32_t* d_status = nullptr; int32_t* h_status = nullptr;
    bool compact = true;                 // lean lattices: persistent tiles over active particles
    int64_t turns_per_launch = 0;        // 0: from the work size
    // one tile queue: a job's sets share one lattice, and with a queue per
    // XCD the waves of the first XCD would take every tile of a small set
    int queues = 1;
    int ppl = 0;                         // particles per lane of the lean until path (0: auto)
    bool mix = true;                     // with ppl auto: mixed one/two-particle launches
    bool balance = true;                 // one tile per wave and launch, furthest behind first
    bool wave_prio = true;               // waves with more turns left issue first
    int staging = 0;                     // lean until path: lattice via scalar loads or LDS
    int64_t max_rec_words = 0;           // longest record of the packed lattice
    int waves_per_cu = 0;                // resident waves per CU (0: the kernel's occupancy)
    st_batch::BatchRunner runner;        // scheduling of the lean until path
    bool runner_ready = false;
    // NS(TrackJob_track_until_noalloc): a runner on the device's scratch
    // memory, with its pinned host arrays made at set-up
    st_batch::BatchRunner na_runner;
    int32_t* na_host = nullptr;
    unsigned char* na_stage = nullptr;
    int n_cu = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    float last_kernel_ms = 0.0f;
    std::vector< st_dev_particles_t > views;
    st_dev_particles_t ebe_view{};
    // introspection arrays (track_job.h:360-470): particles per set,
    // element indices of the beam monitors, the element-by-element
    // configuration as the host sees it
    std::vector< uint64_t > set_sizes_u64, monitor_indices_u64;
    st_ebe_config_t ebe_cfg_host{};
    // particle addresses (track_job.h:206-270): one NS(ParticlesAddr) per
    // Particles object of the particle buffer, holding the device
    // addresses of the job's copy of its 21 field arrays
    NS(Buffer)* paddr_buf = nullptr;
    bool has_paddr = false;
    // distinct (destination, source) pairs of the assignment items, and
    // the items of one pair as CObjects (rebuilt on request)
    std::vector< NS(AssignAddressItemDestSrcPair) > pair_cache;
    NS(Buffer)* item_objs = nullptr;
    uint64_t items_version = 0, item_objs_version = ~0ull, item_objs_dest = 0, item_objs_src = 0;

    ~NS(TrackJobBase)()
    {
        release();
        for( Stored& b : stored ) if( b.owned && b.buf ) NS(Buffer_delete)( b.buf );
        if( item_objs ) NS(Buffer_delete)( item_objs );
    }

    // -- stored buffers and address assignments (track_job_base.cpp:
    // 1133-1260, 2748-2805): buffer ids 1 particles, 2 beam elements,
    // 3 output, 1024 + k the k-th stored buffer
    struct Stored { NS(Buffer)* buf = nullptr; bool owned = false; };
    std::vector< Stored > stored;
    std::vector< NS(AssignAddressItem) > assign_items;

    NS(Buffer)* buffer_by_id( uint64_t id ) const
    {
        if( id == ST_ARCH_PARTICLES_BUFFER_ID ) return pbuf;
        if( id == ST_ARCH_BEAM_ELEMENTS_BUFFER_ID ) return ebuf;
        if( id == ST_ARCH_OUTPUT_BUFFER_ID ) return obuf;
        if( id >= ST_ARCH_MIN_USER_DEFINED_BUFFER_ID &&
            id - ST_ARCH_MIN_USER_DEFINED_BUFFER_ID < stored.size() )
            return stored[ id - ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ].buf;
        return nullptr;
    }

    // address of field ``offset`` of object ``index`` of type ``type`` in
    // ``b`` (host address), 0 when there is no such object / field
    static uint64_t field_addr( NS(Buffer)* b, uint64_t type, uint64_t index, uint64_t offset,
                                uint64_t width )
    {
        if( b == nullptr || index >= NS(Buffer_get_num_of_objects)( b ) ) return 0u;
        NS(Object) const* o = NS(Buffer_get_const_object)( b, index );
        if( o == nullptr || o->type_id != type || offset + width > o->size ) return 0u;
        return o->begin_addr + offset;
    }

    // performs the items with the given (dest, src) pair, or all of them;
    // the device lattice is rebuilt when a beam element was changed
    bool assign( bool all, uint64_t dest_id, uint64_t src_id )
    {
        bool lattice_changed = false;
        for( const NS(AssignAddressItem)& it : assign_items )
        {
            if( !all && ( it.dest_buffer_id != dest_id || it.src_buffer_id != src_id ) ) continue;
            if( it.dest_buffer_id != ST_ARCH_BEAM_ELEMENTS_BUFFER_ID )
            { g_err = "address assignments into the particle or output buffer are not supported"; return false; }
            uint64_t const dst = field_addr( buffer_by_id( it.dest_buffer_id ), it.dest_elem_type_id,
                                             it.dest_elem_index, it.dest_pointer_offset, 8u );
            uint64_t const src = field_addr( buffer_by_id( it.src_buffer_id ), it.src_elem_type_id,
                                             it.src_elem_index, it.src_pointer_offset, 0u );
            if( dst == 0u || src == 0u || ( dst % 8u ) != 0u )
            { g_err = "address assignment item does not name existing objects"; return false; }
            std::memcpy( reinterpret_cast< void* >( static_cast< uintptr_t >( dst ) ), &src, 8 );
            lattice_changed = true;
        }
        return !lattice_changed || d_lat == nullptr || pack_lattice();
    }

    void release()
    {
        if( d_pbuf ) ( void )hipFree( d_pbuf );
        if( d_obuf ) ( void )hipFree( d_obuf );
        if( d_lat ) ( void )hipFree( d_lat );
        if( d_status ) ( void )hipFree( d_status );
        runner.release();
        runner_ready = false;
        na_runner.release();
        if( na_host ) ( void )hipHostFree( na_host );
        if( na_stage ) ( void )hipHostFree( na_stage );
        na_host = nullptr; na_stage = nullptr;
        if( h_status ) ( void )hipHostFree( h_status );
        if( ev_start ) ( void )hipEventDestroy( ev_start );
        if( ev_stop ) ( void )hipEventDestroy( ev_stop );
        if( stream ) ( void )hipStreamDestroy( stream );
        d_pbuf = d_obuf = nullptr; d_lat = nullptr; d_status = nullptr;
        h_status = nullptr; stream = nullptr; ev_start = ev_stop = nullptr;
        if( owns_obuf && obuf ) NS(Buffer_delete)( obuf );
        obuf = nullptr; owns_obuf = false;
        if( paddr_buf ) NS(Buffer_delete)( paddr_buf );
        paddr_buf = nullptr; has_paddr = false;
    }

    // NS(ParticlesAddr_prepare_buffer_based_on_particles_buffer): one
    // zeroed address record per Particles object of the particle buffer
    bool prepare_paddr()
    {
        paddr_buf = NS(Buffer_new)( 0 );
        if( paddr_buf == nullptr ) { g_err = "cannot allocate the particle-address buffer"; return false; }
        uint64_t const n = NS(Buffer_get_num_of_objects)( pbuf );
        for( uint64_t k = 0; k < n; ++k )
        {
            NS(ParticlesAddr) a{};
            if( NS(Buffer_add_object)( paddr_buf, &a, sizeof( a ), ST_OBJECT_TYPE_PARTICLES_ADDR, 0,
                                       nullptr, nullptr, nullptr ) == nullptr )
            { g_err = "cannot fill the particle-address buffer"; return false; }
        }
        has_paddr = false;
        return true;
    }

    NS(ParticlesAddr)* paddr( uint64_t k ) const
    {
        if( paddr_buf == nullptr || k >= NS(Buffer_get_num_of_objects)( paddr_buf ) ) return nullptr;
        NS(Object)* o = NS(Buffer_get_object)( paddr_buf, k );
        return o ? ( NS(ParticlesAddr)* )( uintptr_t )o->begin_addr : nullptr;
    }

    // the device addresses of every Particles object of the job's device
    // copy of the particle buffer (the reference's kernel fetches them on
    // the device; here they follow from the image's relocation)
    bool fetch_paddr()
    {
        if( paddr_buf == nullptr || d_pbuf == nullptr ) return false;
        uint64_t const n = NS(Buffer_get_num_of_objects)( pbuf );
        for( uint64_t k = 0; k < n; ++k )
        {
            NS(ParticlesAddr)* a = paddr( k );
            if( a == nullptr ) return false;
            std::memset( a, 0, sizeof( *a ) );
            if( NS(Particles_buffer_get_const_particles)( pbuf, k ) == nullptr ) continue;
            st_dev_particles_t const v = device_view( pbuf, d_pbuf, k );
            a->num_particles = v.n;
            uint64_t* dst = &a->q0_addr;
            double* const* fr = &v.q0;
            for( int f = 0; f < 17; ++f ) dst[ f ] = ( uint64_t )( uintptr_t )fr[ f ];
            int64_t* const* fi = &v.particle_id;
            for( int f = 0; f < 4; ++f ) dst[ 17 + f ] = ( uint64_t )( uintptr_t )fi[ f ];
        }
        has_paddr = true;
        return true;
    }

    // -- host <-> device images ------------------------------------------
    bool upload_image( NS(Buffer)* b, unsigned char* d, size_t bytes )
    {
        std::vector< uint64_t > staging( ( bytes + 7 ) / 8 );
        std::memcpy( staging.data(), NS(Buffer_get_const_data_begin)( b ), bytes );
        if( NS(ManagedBuffer_remap_to)( reinterpret_cast< unsigned char* >( staging.data() ),
                                        ( NS(buffer_addr_t) )( uintptr_t )d ) != 0 )
        { g_err = "relocation of buffer image failed"; return false; }
        return hip_ok( hipMemcpyAsync( d, staging.data(), bytes, hipMemcpyHostToDevice, stream ),
                       "hipMemcpyAsync H2D" ) &&
               hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" );
    }

    bool download_image( NS(Buffer)* b, const unsigned char* d, size_t bytes )
    {
        unsigned char* h = NS(Buffer_get_data_begin)( b );
        if( NS(Buffer_get_size)( b ) != bytes ) { g_err = "host buffer changed size"; return false; }
        if( !hip_ok( hipMemcpyAsync( h, d, bytes, hipMemcpyDeviceToHost, stream ), "hipMemcpyAsync D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) ) return false;
        return NS(ManagedBuffer_remap_to)( h, 0u ) == 0;
    }

    st_dev_particles_t device_view( NS(Buffer)* host, unsigned char* d, uint64_t index )
    {
        st_dev_particles_t v{};
        NS(Particles)* hp = NS(Particles_buffer_get_particles)( host, index );
        uint64_t const hb = NS(Buffer_get_data_begin_addr)( host );
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

    bool pack_lattice()
    {
        runner_ready = false;                   // seeds point at the lattice
        int64_t n_elem_l = 0, n_mon = 0;
        const unsigned char* eb = NS(Buffer_get_const_data_begin)( ebuf );
        // stored buffers may hold line-density tables of space-charge elements
        std::vector< const unsigned char* > tables;
        for( const Stored& b : stored )
            if( b.buf ) tables.push_back( NS(Buffer_get_const_data_begin)( b.buf ) );
        int64_t const nt = ( int64_t )tables.size();
        int64_t const nw = NS(Lattice_pack_with_tables)( eb, tables.data(), nt, nullptr, 0, nullptr,
                                                         &n_elem_l, nullptr, &n_mon );
        if( nw < 0 ) { g_err = "lattice could not be packed"; return false; }
        // zero padding: the kernels read the header one record ahead and
        // fetch multipole coefficients in batches of up to 16 words
        lat_host.assign( nw + ST_LAT_PAD_WORDS, 0u );
        rec_offsets.assign( n_elem_l + 1, 0u );
        monitor_fields.assign( n_mon, 0u );
        if( NS(Lattice_pack_with_tables)( eb, tables.data(), nt, lat_host.data(), nw,
                                          rec_offsets.data(), &n_elem_l, monitor_fields.data(),
                                          &n_mon ) != nw ) return false;
        n_elem = n_elem_l;
        lat_words = nw;
        lat_full = false;
        max_rec_words = 0;
        for( int64_t i = 0; i < n_elem; ++i )
        {
            unsigned const op = ST_REC_OP( lat_host[ rec_offsets[ i ] ] );
            if( !ST_OP_IS_LEAN( op ) ) lat_full = true;
            max_rec_words = std::max< int64_t >( max_rec_words, rec_offsets[ i + 1 ] - rec_offsets[ i ] );
        }
        // point the monitors that have an output set at its device header
        if( plan.monitor_elems.size() > 0 && d_obuf != nullptr )
        {
            uint64_t const hb = NS(Buffer_get_data_begin_addr)( obuf );
            size_t m = 0;
            for( int64_t i = 0; i < n_elem && m < plan.monitor_elems.size(); ++i )
            {
                if( ST_REC_OP( lat_host[ rec_offsets[ i ] ] ) != ST_OP_MONITOR ) continue;
                uint32_t const field = rec_offsets[ i ] + ST_REC_PAYLOAD + 3u;
                if( i == plan.monitor_elems[ m ] )
                {
                    NS(Object) const* o = NS(Buffer_get_const_object)( obuf, monitor_offset + m );
                    lat_host[ field ] = ( uint64_t )( uintptr_t )d_obuf + ( o->begin_addr - hb );
                    ++m;
                }
            }
        }
        if( d_lat ) { ( void )hipFree( d_lat ); d_lat = nullptr; }
        if( !hip_ok( hipMalloc( &d_lat, lat_host.size() * 8u ), "hipMalloc lattice" ) ) return false;
        return hip_ok( hipMemcpyAsync( d_lat, lat_host.data(), lat_host.size() * 8u,
                                       hipMemcpyHostToDevice, stream ), "upload lattice" ) &&
               hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" );
    }

    bool prepare_output()
    {
        std::vector< int64_t > need;
        if( plan.has_ebe ) need.push_back( plan.ebe_size );
        need.insert( need.end(), plan.monitor_sizes.begin(), plan.monitor_sizes.end() );
        if( need.empty() ) return true;
        if( obuf == nullptr ) { obuf = NS(Buffer_new)( 0 ); owns_obuf = true; }
        if( obuf == nullptr ) { g_err = "cannot allocate output buffer"; return false; }
        // reuse a caller-prepared buffer whose leading sets match exactly
        uint64_t base = NS(Buffer_get_num_of_objects)( obuf );
        bool match = base >= need.size();
        for( size_t k = 0; match && k < need.size(); ++k )
        {
            NS(Particles) const* p = NS(Particles_buffer_get_const_particles)( obuf, k );
            match = p != nullptr && p->num_particles == need[ k ];
        }
        if( match ) base = 0;
        else
            for( int64_t sz : need )
                if( NS(Particles_new)( obuf, sz ) == nullptr )
                { g_err = "output buffer cannot hold the required particle sets"; return false; }
        ebe_offset = base;
        monitor_offset = base + ( plan.has_ebe ? 1u : 0u );
        // be_monitor/output_buffer.h:96-140: out_address and particle range
        for( size_t m = 0; m < plan.monitor_elems.size(); ++m )
        {
            NS(Object)* eo = NS(Buffer_get_object)( ebuf, plan.monitor_elems[ m ] );
            NS(BeamMonitor)* mon = ( NS(BeamMonitor)* )( uintptr_t )eo->begin_addr;
            mon->min_particle_id = attr.min_pid;
            mon->max_particle_id = attr.max_pid;
            mon->out_address = NS(Buffer_get_const_object)( obuf, monitor_offset + m )->begin_addr;
        }
        if( plan.has_ebe )
        {
            st_ebe_config_t& c = ebe_cfg;
            c.order = 0;
            c.num_particles_to_store = attr.max_pid - attr.min_pid + 1;
            c.num_elements_to_store = attr.max_elem - attr.min_elem + 1;
            c.num_turns_to_store = plan.max_ebe_turn - attr.min_turn + 1;
            c.min_particle_id = attr.min_pid; c.max_particle_id = attr.max_pid;
            c.min_element_id = attr.min_elem; c.max_element_id = attr.max_elem;
            c.min_turn = attr.min_turn; c.max_turn = plan.max_ebe_turn;
            c.is_rolling = ebe_rolling ? 1 : 0;
            c.out_store_addr = NS(Buffer_get_const_object)( obuf, ebe_offset )->begin_addr;
        }
        return true;
    }

    bool setup( NS(Buffer)* p, uint64_t nsets, const uint64_t* idx, NS(Buffer)* e,
                NS(Buffer)* o, uint64_t ute )
    {
        release();
        if( p == nullptr || e == nullptr ) { g_err = "particles and beam elements buffers required"; return false; }
        if( NS(Buffer_needs_remapping)( p ) || NS(Buffer_needs_remapping)( e ) ||
            ( o != nullptr && NS(Buffer_needs_remapping)( o ) ) )
        { g_err = "buffers must not need remapping"; return false; }
        pbuf = p; ebuf = e; obuf = o; owns_obuf = false; until_turn_ebe = ute;
        sets.clear(); set_sizes.clear();
        if( nsets == 0 || idx == nullptr ) sets.push_back( 0u );
        else sets.assign( idx, idx + nsets );
        for( uint64_t k : sets )
        {
            NS(Particles) const* pp = NS(Particles_buffer_get_const_particles)( pbuf, k );
            if( pp == nullptr ) { g_err = "particle set index does not name a Particles object"; return false; }
            set_sizes.push_back( pp->num_particles );
        }
        attr = Attr();
        plan = OutputPlan();
        if( !find_attributes( pbuf, sets, ebuf, attr, nullptr ) ) return false;
        plan_output( ebuf, attr, until_turn_ebe, ebe_passes, plan );
        if( !prepare_output() ) return false;

        if( !hip_ok( hipSetDevice( device ), "hipSetDevice" ) ) return false;
        if( !hip_ok( hipStreamCreate( &stream ), "hipStreamCreate" ) ) return false;
        if( !hip_ok( hipEventCreate( &ev_start ), "hipEventCreate" ) ||
            !hip_ok( hipEventCreate( &ev_stop ), "hipEventCreate" ) ) return false;
        if( !hip_ok( hipMalloc( &d_status, sizeof( int32_t ) ), "hipMalloc status" ) ) return false;
        if( !hip_ok( hipHostMalloc( &h_status, sizeof( int32_t ) ), "hipHostMalloc" ) ) return false;
        if( !hip_ok( hipDeviceGetAttribute( &n_cu, hipDeviceAttributeMultiprocessorCount, device ),
                     "hipDeviceGetAttribute" ) ) return false;
        // what the no-allocation entry point needs besides the device's
        // scratch memory, made here so that it allocates nothing itself
        if( ( int64_t )sets.size() <= ST_NOALLOC_MAX_SETS )
        {
            int64_t const S = ( int64_t )sets.size();
            na_runner.n_part.reserve( S );
            na_runner.n_elem.reserve( S );
            if( !hip_ok( hipHostMalloc( &na_host, st_batch::BatchRunner::host_ints( S ) * sizeof( int32_t ) ),
                         "hipHostMalloc" ) ||
                !hip_ok( hipHostMalloc( &na_stage, st_batch::BatchRunner::stage_bytes( S ) ),
                         "hipHostMalloc" ) ) return false;
            if( scratch_address( device, true ) == nullptr ) return false;
        }
        pbuf_bytes = NS(Buffer_get_size)( pbuf );
        if( !hip_ok( hipMalloc( &d_pbuf, pbuf_bytes ), "hipMalloc particles" ) ) return false;
        if( !upload_image( pbuf, d_pbuf, pbuf_bytes ) ) return false;
        if( obuf != nullptr )
        {
            obuf_bytes = NS(Buffer_get_size)( obuf );
            if( !hip_ok( hipMalloc( &d_obuf, obuf_bytes ), "hipMalloc output" ) ) return false;
            if( !upload_image( obuf, d_obuf, obuf_bytes ) ) return false;
        }
        if( !pack_lattice() ) return false;
        views.clear();
        for( uint64_t k : sets ) views.push_back( device_view( pbuf, d_pbuf, k ) );
        if( plan.has_ebe ) ebe_view = device_view( obuf, d_obuf, ebe_offset );
        set_sizes_u64.assign( set_sizes.begin(), set_sizes.end() );
        monitor_indices_u64.assign( plan.monitor_elems.begin(), plan.monitor_elems.end() );
        ebe_cfg_host = ebe_cfg;
        return prepare_paddr();
    }

    // status word cleared before and read back after the launches; the two
    // events bracket the kernels on the job stream (kernel time only)
    bool begin_status()
    {
        return hip_ok( hipMemsetAsync( d_status, 0, sizeof( int32_t ), stream ), "hipMemsetAsync" ) &&
               hip_ok( hipEventRecord( ev_start, stream ), "hipEventRecord" );
    }

    int32_t end_status( bool launched_ok )
    {
        if( !launched_ok ) { last_status = -1; return -1; }
        if( !hip_ok( hipEventRecord( ev_stop, stream ), "hipEventRecord" ) ||
            !hip_ok( hipMemcpyAsync( h_status, d_status, sizeof( int32_t ),
                                     hipMemcpyDeviceToHost, stream ), "status D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) )
        { last_status = -1; return -1; }
        float ms = 0.0f;
        last_kernel_ms = ( hipEventElapsedTime( &ms, ev_start, ev_stop ) == hipSuccess ) ? ms : -1.0f;
        last_status = *h_status;
        return last_status;
    }

    // track_until (the default): the persistent scheduling of the batched
    // form with the job's particle sets as seeds (batch_runner.h).  Only
    // the particles with turns to go take lanes; a long request is split
    // into chunks of a few turns (list rebuilt between them, on the
    // device), so lanes of particles lost during a chunk are reused at the
    // next.  Every particle is tracked turn by turn on its own, so the
    // results are those of one launch.  Lattices with maps outside the
    // lean set run the heavy variant of the same loop (those records out
    // of line).  By default (balance=1) each launch gives every resident
    // wave one tile, the particles furthest behind first, and the host
    // reads the counts back between launches; balance=0 hands every wave
    // tiles until none are left in each chunk.  compact=0 runs one launch
    // with one lane per particle.
    st_batch::Config runner_config() const
    {
        st_batch::Config c;
        c.fma = fma;
        c.block = ( block == 64 || block == 128 ) ? block : 256;
        c.turns_per_launch = turns_per_launch;
        c.queues = queues;
        c.heavy = lat_full;
        c.ppl = ( lat_full || ppl != 2 ) ? 1 : 2;
        // auto: two particles per lane for whole requests only in the FMA
        // build (the exact build's longer dependency chains leave nothing
        // for the pair to cover: the same rate with every wave busy, 5 %
        // slower over the 1000-turn job of the 20,000-element ring); the
        // mixed launches, which keep every SIMD equally loaded, in both
        // (2.6 % faster over that job in the exact build)
        c.ppl_auto = !lat_full && ppl == 0 && staging == 0 && fma;
        c.mix = mix && !lat_full && ppl == 0 && staging == 0;
        c.balance = balance;
        c.wave_prio = wave_prio;
        c.staging = ( lat_full || c.ppl == 2 ) ? 0 : staging;
        // resident waves per CU: 8 per SIMD (the lean kernels' register
        // budget, with one particle per lane and with two)
        c.waves_per_cu = 32;
        if( waves_per_cu > 0 ) c.waves_per_cu = std::min( waves_per_cu, c.waves_per_cu );
        c.graph = false;
        return c;
    }

    bool ensure_runner()
    {
        if( runner_ready ) return true;
        runner.cfg = runner_config();
        runner.max_record_words = max_rec_words;
        std::vector< st_batch_seed_t > seeds;
        for( const st_dev_particles_t& v : views )
        {
            st_batch_seed_t sd{};
            sd.P = v; sd.lat = d_lat; sd.n_elem = n_elem; sd.list = nullptr;
            seeds.push_back( sd );
        }
        runner_ready = runner.init( seeds, stream );
        return runner_ready;
    }

    int32_t track_until( int64_t until_turn )
    {
        if( !begin_status() ) return end_status( false );
        bool ok = true;
        if( compact && n_elem > 0 )
        {
            int64_t n_active = 0, min_turn = 0;
            if( balance ) return end_status( ensure_runner() &&
                                             runner.run_balanced( until_turn, d_status, stream ) );
            ok = ensure_runner() && runner.select_sync( until_turn, stream, n_active, min_turn );
            if( ok && n_active > 0 )
            {
                int64_t const k = runner.chunk_turns( n_active );
                int64_t const chunks = ( until_turn - min_turn + k - 1 ) / k;
                ok = runner.run( until_turn, chunks, k, d_status, stream );
            }
            return end_status( ok );
        }
        for( const st_dev_particles_t& v : views )
        {
            hipError_t const err = fma
                ? st_launch_track_until_fma( &v, d_lat, n_elem, until_turn, lat_full, d_status, block, stream )
                : st_launch_track_until_exact( &v, d_lat, n_elem, until_turn, lat_full, d_status, block, stream );
            ok = ok && hip_ok( err, "launch track_until" );
        }
        return end_status( ok );
    }

    // track_until with every scratch array of the call in the device's
    // ``scratch`` memory: the same scheduling and kernels, nothing
    // allocated or freed.  Calls from several jobs on one device share that
    // memory; calls that overlap in time on one device race.
    int32_t track_until_noalloc( int64_t until_turn )
    {
        int64_t const S = ( int64_t )views.size();
        int64_t total = 0;
        for( const st_dev_particles_t& v : views ) total += v.n;
        unsigned char* const base = ( unsigned char* )scratch_address( device, false );
        if( base == nullptr || na_host == nullptr || na_stage == nullptr )
        { g_err = "no-allocation tracking: the device's scratch memory is not set up"; last_status = -1; return -1; }
        if( S > ST_NOALLOC_MAX_SETS || total > ST_NOALLOC_MAX_PARTICLES )
        {
            char m[ 160 ];
            std::snprintf( m, sizeof( m ), "no-allocation tracking takes at most %lld particle sets "
                           "of together %lld particles", ( long long )ST_NOALLOC_MAX_SETS,
                           ( long long )ST_NOALLOC_MAX_PARTICLES );
            g_err = m;
            last_status = -1;
            return -1;
        }
        int32_t* const status = ( int32_t* )base;
        if( !hip_ok( hipMemsetAsync( status, 0, sizeof( int32_t ), stream ), "hipMemsetAsync" ) ||
            !hip_ok( hipEventRecord( ev_start, stream ), "hipEventRecord" ) )
        { last_status = -1; return -1; }
        bool ok = true;
        if( compact && n_elem > 0 )
        {
            na_runner.cfg = runner_config();
            na_runner.max_record_words = max_rec_words;
            ok = na_runner.init_borrowed( views.data(), d_lat, n_elem, S,
                                          base + ST_NOALLOC_STATUS_BYTES,
                                          ST_NOALLOC_SCRATCH_BYTES - ST_NOALLOC_STATUS_BYTES,
                                          na_host, na_stage, n_cu, stream );
            if( ok && balance ) ok = na_runner.run_balanced( until_turn, status, stream );
            else if( ok )
            {
                int64_t n_active = 0, min_turn = 0;
                ok = na_runner.select_sync( until_turn, stream, n_active, min_turn );
                if( ok && n_active > 0 )
                {
                    int64_t const k = na_runner.chunk_turns( n_active );
                    ok = na_runner.run( until_turn, ( until_turn - min_turn + k - 1 ) / k, k, status,
                                        stream );
                }
            }
        }
        else
            for( const st_dev_particles_t& v : views )
            {
                hipError_t const err = fma
                    ? st_launch_track_until_fma( &v, d_lat, n_elem, until_turn, lat_full, status, block, stream )
                    : st_launch_track_until_exact( &v, d_lat, n_elem, until_turn, lat_full, status, block, stream );
                ok = ok && hip_ok( err, "launch track_until" );
            }
        if( !ok ||
            !hip_ok( hipEventRecord( ev_stop, stream ), "hipEventRecord" ) ||
            !hip_ok( hipMemcpyAsync( h_status, status, sizeof( int32_t ), hipMemcpyDeviceToHost, stream ),
                     "status D2H" ) ||
            !hip_ok( hipStreamSynchronize( stream ), "hipStreamSynchronize" ) )
        {
            ( void )hipStreamSynchronize( stream );
            last_status = -1;
            return -1;
        }
        float ms = 0.0f;
        last_kernel_ms = ( hipEventElapsedTime( &ms, ev_start, ev_stop ) == hipSuccess ) ? ms : -1.0f;
        last_status = *h_status;
        return last_status;
    }

    int32_t track_ebe( int64_t until_turn )
    {
        if( !plan.has_ebe ) return track_until( until_turn );
        if( !begin_status() ) return end_status( false );
        bool ok = true;
        // the smallest at_turn of the live particles (one select pass, read
        // back) sets how many one-turn windows the launcher runs
        int64_t from_turn = -1;
        if( compact && n_elem > 0 )
        {
            int64_t n_active = 0, min_turn = 0;
            if( !ensure_runner() || !runner.select_sync( until_turn, stream, n_active, min_turn ) )
                return end_status( false );
            from_turn = n_active > 0 ? min_turn : until_turn;
        }
        for( const st_dev_particles_t& v : views )
        {
            hipError_t const err = fma
                ? st_launch_track_ebe_fma( &v, d_lat, n_elem, until_turn, &ebe_device_cfg(), plan.ebe_cap, &ebe_view, d_status, from_turn, block, stream )
                : st_launch_track_ebe_exact( &v, d_lat, n_elem, until_turn, &ebe_device_cfg(), plan.ebe_cap, &ebe_view, d_status, from_turn, block, stream );
            ok = ok && hip_ok( err, "launch track_elem_by_elem" );
        }
        return end_status( ok );
    }

    st_ebe_config_t dev_cfg{};
    const st_ebe_config_t& ebe_device_cfg()
    {
        dev_cfg = ebe_cfg;
        dev_cfg.out_store_addr = ( uint64_t )( uintptr_t )d_obuf +
            ( ebe_cfg.out_store_addr - NS(Buffer_get_data_begin_addr)( obuf ) );
        return dev_cfg;
    }

    int32_t track_line( uint64_t b, uint64_t e, bool finish )
    {
        if( b > e || e > ( uint64_t )n_elem ) { last_status = -1; return -1; }
        if( !begin_status() ) return end_status( false );
        bool ok = true;
        const uint64_t* lat = d_lat + rec_offsets[ b ];
        for( const st_dev_particles_t& v : views )
        {
            hipError_t const err = fma
                ? st_launch_track_line_fma( &v, lat, ( int64_t )( e - b ), finish, d_status, block, stream )
                : st_launch_track_line_exact( &v, lat, ( int64_t )( e - b ), finish, d_status, block, stream );
            ok = ok && hip_ok( err, "launch track_line" );
        }
        return end_status( ok );
    }

    void parse_config( const char* cfg )
    {
        if( cfg == nullptr ) return;
        config = cfg;
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
            else if( k == "block" ) block = ( int )iv;
            else if( k == "device" ) { device = ( int )iv; device_id = v; }
            else if( k == "ebe_rolling" ) ebe_rolling = iv != 0;
            else if( k == "ebe_store_passes" ) ebe_passes = iv > 0 ? iv : 0;
            else if( k == "compact" ) compact = iv != 0;
            else if( k == "turns_per_launch" ) turns_per_launch = iv > 0 ? iv : 0;
            else if( k == "queues" ) queues = ( iv == 8 ) ? 8 : 1;
            else if( k == "ppl" ) ppl = ( v == "auto" ) ? 0 : ( iv == 2 ) ? 2 : 1;
            else if( k == "mix" ) mix = iv != 0;
            else if( k == "balance" ) balance = iv != 0;
            else if( k == "wave_prio" ) wave_prio = iv != 0;
            else if( k == "staging" ) staging = ( v == "lds" || v == "1" ) ? 1 : 0;
            else if( k == "waves_per_cu" ) waves_per_cu = ( iv >= 1 && iv <= 32 ) ? ( int )iv : 0;
        }
    }
};

namespace
{
bool parse_arch( const char* arch, NS(TrackJobBase)* job )
{
    if( arch == nullptr ) return false;
    std::string a = arch;
    for( char& c : a ) c = ( char )std::tolower( ( unsigned char )c );
    size_t const colon = a.find( ':' );
    if( colon != std::string::npos )
    {
        job->device_id = a.substr( colon + 1 );
        job->device = ( int )std::strtol( job->device_id.c_str(), nullptr, 10 );
        a = a.substr( 0, colon );
    }
    return a == "hip" || a == "amd" || a == "amdgpu" || a == "gpu";
}

NS(TrackJobBase)* make_job( const char* arch, const char* config )
{
    NS(TrackJobBase)* job = new NS(TrackJobBase)();
    if( !parse_arch( arch, job ) )
    {
        g_err = std::string( "unsupported architecture '" ) + ( arch ? arch : "(null)" ) +
                "' (this library provides \"hip\" only)";
        delete job;
        return nullptr;
    }
    job->parse_config( config );
    return job;
}
}  // namespace

extern "C" {

NS(TrackJobBase)* NS(TrackJob_create)( char const* arch, char const* config_str )
{
    return make_job( arch, config_str );
}

NS(TrackJobBase)* NS(TrackJob_new_detailed)( char const* arch, NS(Buffer)* pbuf,
    NS(buffer_size_t) nsets, NS(buffer_size_t) const* idx, NS(Buffer)* ebuf,
    NS(Buffer)* obuf, NS(buffer_size_t) until_turn_ebe, char const* config_str )
{
    NS(TrackJobBase)* job = make_job( arch, config_str );
    if( job == nullptr ) return nullptr;
    if( !job->setup( pbuf, nsets, idx, ebuf, obuf, until_turn_ebe ) ) { delete job; return nullptr; }
    return job;
}

NS(TrackJobBase)* NS(TrackJob_new_with_output)( char const* arch, NS(Buffer)* pbuf,
    NS(Buffer)* ebuf, NS(Buffer)* obuf, NS(buffer_size_t) until_turn_ebe,
    char const* config_str )
{
    return NS(TrackJob_new_detailed)( arch, pbuf, 0u, nullptr, ebuf, obuf, until_turn_ebe, config_str );
}

NS(TrackJobBase)* NS(TrackJob_new)( char const* arch, NS(Buffer)* pbuf,
    NS(Buffer)* ebuf, char const* config_str )
{
    return NS(TrackJob_new_detailed)( arch, pbuf, 0u, nullptr, ebuf, nullptr, 0u, config_str );
}

void NS(TrackJob_delete)( NS(TrackJobBase)* job ) { delete job; }

NS(track_status_t) NS(TrackJob_track_until)( NS(TrackJobBase)* job, NS(buffer_size_t) until_turn )
{
    return ( job && job->d_pbuf ) ? job->track_until( ( int64_t )until_turn ) : -1;
}

NS(track_status_t) NS(TrackJob_track_until_noalloc)( NS(TrackJobBase)* job, NS(buffer_size_t) until_turn )
{
    return ( job && job->d_pbuf ) ? job->track_until_noalloc( ( int64_t )until_turn ) : -1;
}

NS(track_status_t) NS(TrackJob_track_elem_by_elem)( NS(TrackJobBase)* job, NS(buffer_size_t) until_turn )
{
    return ( job && job->d_pbuf ) ? job->track_ebe( ( int64_t )until_turn ) : -1;
}

NS(track_status_t) NS(TrackJob_track_line)( NS(TrackJobBase)* job, NS(buffer_size_t) b,
                                            NS(buffer_size_t) e, bool finish_turn )
{
    return ( job && job->d_pbuf ) ? job->track_line( b, e, finish_turn ) : -1;
}

void NS(TrackJob_collect_detailed)( NS(TrackJobBase)* job, NS(track_job_collect_flag_t) flags )
{
    if( job == nullptr || job->d_pbuf == nullptr ) return;
    if( flags & ST_TRACK_JOB_IO_PARTICLES )
        job->download_image( job->pbuf, job->d_pbuf, job->pbuf_bytes );
    if( ( flags & ST_TRACK_JOB_IO_OUTPUT ) && job->obuf && job->d_obuf )
        job->download_image( job->obuf, job->d_obuf, job->obuf_bytes );
    /* beam elements are never modified on the device: nothing to collect */
}

void NS(TrackJob_collect)( NS(TrackJobBase)* job )
{
    if( job ) NS(TrackJob_collect_detailed)( job, job->collect_flags );
}
void NS(TrackJob_collect_particles)( NS(TrackJobBase)* job ) { NS(TrackJob_collect_detailed)( job, ST_TRACK_JOB_IO_PARTICLES ); }
void NS(TrackJob_collect_beam_elements)( NS(TrackJobBase)* job ) { NS(TrackJob_collect_detailed)( job, ST_TRACK_JOB_IO_BEAM_ELEMENTS ); }
void NS(TrackJob_collect_output)( NS(TrackJobBase)* job ) { NS(TrackJob_collect_detailed)( job, ST_TRACK_JOB_IO_OUTPUT ); }

static void set_flag( NS(TrackJobBase)* job, uint16_t f, bool on )
{
    if( job ) job->collect_flags = on ? ( job->collect_flags | f ) : ( job->collect_flags & ~f );
}
void NS(TrackJob_enable_collect_particles)( NS(TrackJobBase)* j ) { set_flag( j, ST_TRACK_JOB_IO_PARTICLES, true ); }
void NS(TrackJob_disable_collect_particles)( NS(TrackJobBase)* j ) { set_flag( j, ST_TRACK_JOB_IO_PARTICLES, false ); }
bool NS(TrackJob_is_collecting_particles)( const NS(TrackJobBase)* j ) { return j && ( j->collect_flags & ST_TRACK_JOB_IO_PARTICLES ); }
void NS(TrackJob_enable_collect_beam_elements)( NS(TrackJobBase)* j ) { set_flag( j, ST_TRACK_JOB_IO_BEAM_ELEMENTS, true ); }
void NS(TrackJob_disable_collect_beam_elements)( NS(TrackJobBase)* j ) { set_flag( j, ST_TRACK_JOB_IO_BEAM_ELEMENTS, false ); }
bool NS(TrackJob_is_collecting_beam_elements)( const NS(TrackJobBase)* j ) { return j && ( j->collect_flags & ST_TRACK_JOB_IO_BEAM_ELEMENTS ); }
void NS(TrackJob_enable_collect_output)( NS(TrackJobBase)* j ) { set_flag( j, ST_TRACK_JOB_IO_OUTPUT, true ); }
void NS(TrackJob_disable_collect_output)( NS(TrackJobBase)* j ) { set_flag( j, ST_TRACK_JOB_IO_OUTPUT, false ); }
bool NS(TrackJob_is_collecting_output)( const NS(TrackJobBase)* j ) { return j && ( j->collect_flags & ST_TRACK_JOB_IO_OUTPUT ); }
NS(track_job_collect_flag_t) NS(TrackJob_get_collect_flags)( const NS(TrackJobBase)* j ) { return j ? j->collect_flags : 0; }
void NS(TrackJob_set_collect_flags)( NS(TrackJobBase)* j, NS(track_job_collect_flag_t) f ) { if( j ) j->collect_flags = f; }
bool NS(TrackJob_requires_collecting)( const NS(TrackJobBase)* j ) { return j != nullptr; }

void NS(TrackJob_push)( NS(TrackJobBase)* job, NS(track_job_push_flag_t) flags )
{
    if( job == nullptr || job->d_pbuf == nullptr ) return;
    if( flags & ST_TRACK_JOB_IO_PARTICLES ) job->upload_image( job->pbuf, job->d_pbuf, job->pbuf_bytes );
    if( flags & ST_TRACK_JOB_IO_BEAM_ELEMENTS ) job->pack_lattice();
    if( ( flags & ST_TRACK_JOB_IO_OUTPUT ) && job->obuf && job->d_obuf )
        job->upload_image( job->obuf, job->d_obuf, job->obuf_bytes );
}
void NS(TrackJob_push_particles)( NS(TrackJobBase)* j ) { NS(TrackJob_push)( j, ST_TRACK_JOB_IO_PARTICLES ); }
void NS(TrackJob_push_beam_elements)( NS(TrackJobBase)* j ) { NS(TrackJob_push)( j, ST_TRACK_JOB_IO_BEAM_ELEMENTS ); }
void NS(TrackJob_push_output)( NS(TrackJobBase)* j ) { NS(TrackJob_push)( j, ST_TRACK_JOB_IO_OUTPUT ); }

void NS(TrackJob_clear)( NS(TrackJobBase)* job ) { if( job ) job->release(); }

bool NS(TrackJob_reset_detailed)( NS(TrackJobBase)* job, NS(Buffer)* pbuf,
    NS(buffer_size_t) nsets, NS(buffer_size_t) const* idx, NS(Buffer)* ebuf,
    NS(Buffer)* obuf, NS(buffer_size_t) ute )
{
    return job != nullptr && job->setup( pbuf, nsets, idx, ebuf, obuf, ute );
}
bool NS(TrackJob_reset_with_output)( NS(TrackJobBase)* job, NS(Buffer)* pbuf,
    NS(Buffer)* ebuf, NS(Buffer)* obuf, NS(buffer_size_t) ute )
{
    return NS(TrackJob_reset_detailed)( job, pbuf, 0u, nullptr, ebuf, obuf, ute );
}
bool NS(TrackJob_reset)( NS(TrackJobBase)* job, NS(Buffer)* pbuf, NS(Buffer)* ebuf, NS(Buffer)* obuf )
{
    return NS(TrackJob_reset_detailed)( job, pbuf, 0u, nullptr, ebuf, obuf, 0u );
}
bool NS(TrackJob_select_particle_set)( NS(TrackJobBase)* job, NS(buffer_size_t) k )
{
    if( job == nullptr ) return false;
    return job->setup( job->pbuf, 1u, &k, job->ebuf, job->owns_obuf ? nullptr : job->obuf,
                       job->until_turn_ebe );
}
bool NS(TrackJob_assign_output_buffer)( NS(TrackJobBase)* job, NS(Buffer)* obuf )
{
    if( job == nullptr ) return false;
    std::vector< uint64_t > s = job->sets;
    return job->setup( job->pbuf, s.size(), s.data(), job->ebuf, obuf, job->until_turn_ebe );
}

NS(track_job_type_t) NS(TrackJob_get_type_id)( const NS(TrackJobBase)* j ) { return j ? 3 : -1; }
char const* NS(TrackJob_get_type_str)( const NS(TrackJobBase)* j ) { return j ? j->arch.c_str() : ""; }
bool NS(TrackJob_has_device_id_str)( const NS(TrackJobBase)* j ) { return j && !j->device_id.empty(); }
char const* NS(TrackJob_get_device_id_str)( const NS(TrackJobBase)* j ) { return j ? j->device_id.c_str() : ""; }
bool NS(TrackJob_has_config_str)( const NS(TrackJobBase)* j ) { return j && !j->config.empty(); }
char const* NS(TrackJob_get_config_str)( const NS(TrackJobBase)* j ) { return j ? j->config.c_str() : ""; }
NS(buffer_size_t) NS(TrackJob_get_num_particle_sets)( const NS(TrackJobBase)* j ) { return j ? j->sets.size() : 0u; }
NS(buffer_size_t) const* NS(TrackJob_get_particle_set_indices_begin)( const NS(TrackJobBase)* j ) { return j ? j->sets.data() : nullptr; }
NS(buffer_size_t) const* NS(TrackJob_get_particle_set_indices_end)( const NS(TrackJobBase)* j ) { return j ? j->sets.data() + j->sets.size() : nullptr; }
NS(buffer_size_t) NS(TrackJob_get_particle_set_index)( const NS(TrackJobBase)* j, NS(buffer_size_t) n ) { return ( j && n < j->sets.size() ) ? j->sets[ n ] : 0u; }
NS(buffer_size_t) NS(TrackJob_get_total_num_particles_in_sets)( const NS(TrackJobBase)* j )
{
    NS(buffer_size_t) t = 0u;
    if( j ) for( int64_t n : j->set_sizes ) t += ( NS(buffer_size_t) )n;
    return t;
}
NS(particle_index_t) NS(TrackJob_get_min_particle_id)( const NS(TrackJobBase)* j ) { return j ? j->attr.min_pid : -1; }
NS(particle_index_t) NS(TrackJob_get_max_particle_id)( const NS(TrackJobBase)* j ) { return j ? j->attr.max_pid : -1; }
NS(particle_index_t) NS(TrackJob_get_min_element_id)( const NS(TrackJobBase)* j ) { return j ? j->attr.min_elem : -1; }
NS(particle_index_t) NS(TrackJob_get_max_element_id)( const NS(TrackJobBase)* j ) { return j ? j->attr.max_elem : -1; }
NS(particle_index_t) NS(TrackJob_get_min_initial_turn_id)( const NS(TrackJobBase)* j ) { return j ? j->attr.min_turn : -1; }
NS(particle_index_t) NS(TrackJob_get_max_initial_turn_id)( const NS(TrackJobBase)* j ) { return j ? j->attr.max_turn : -1; }
NS(Buffer)* NS(TrackJob_get_particles_buffer)( NS(TrackJobBase)* j ) { return j ? j->pbuf : nullptr; }
NS(Buffer)* NS(TrackJob_get_beam_elements_buffer)( NS(TrackJobBase)* j ) { return j ? j->ebuf : nullptr; }
bool NS(TrackJob_has_output_buffer)( const NS(TrackJobBase)* j ) { return j && j->obuf != nullptr; }
bool NS(TrackJob_owns_output_buffer)( const NS(TrackJobBase)* j ) { return j && j->owns_obuf; }
bool NS(TrackJob_has_elem_by_elem_output)( const NS(TrackJobBase)* j ) { return j && j->plan.has_ebe; }
bool NS(TrackJob_has_beam_monitor_output)( const NS(TrackJobBase)* j ) { return j && !j->plan.monitor_elems.empty(); }
NS(buffer_size_t) NS(TrackJob_get_beam_monitor_output_buffer_offset)( const NS(TrackJobBase)* j ) { return j ? j->monitor_offset : 0u; }
NS(buffer_size_t) NS(TrackJob_get_elem_by_elem_output_buffer_offset)( const NS(TrackJobBase)* j ) { return j ? j->ebe_offset : 0u; }
NS(buffer_size_t) NS(TrackJob_get_num_elem_by_elem_turns)( const NS(TrackJobBase)* j ) { return j ? j->until_turn_ebe : 0u; }
NS(buffer_size_t) NS(TrackJob_get_num_beam_monitors)( const NS(TrackJobBase)* j ) { return j ? j->plan.monitor_elems.size() : 0u; }
NS(Buffer)* NS(TrackJob_get_output_buffer)( NS(TrackJobBase)* j ) { return j ? j->obuf : nullptr; }
NS(track_status_t) NS(TrackJob_last_status)( const NS(TrackJobBase)* j ) { return j ? j->last_status : -1; }
double NS(TrackJob_last_kernel_time_ms)( const NS(TrackJobBase)* j ) { return j ? ( double )j->last_kernel_ms : -1.0; }
int NS(TrackJob_uses_fma)( const NS(TrackJobBase)* j ) { return ( j && j->fma ) ? 1 : 0; }
// debug mode of the new-style API (track_job_new.cpp)
extern "C" bool st_internal_job_debug_mode( const NS(TrackJobBase)* j )
{
    return j != nullptr && j->debug_mode;
}

extern "C" void st_internal_job_set_debug_mode( NS(TrackJobBase)* j, bool on )
{
    if( j != nullptr ) j->debug_mode = on;
}

int NS(TrackJob_uses_full_kernel)( const NS(TrackJobBase)* j ) { return ( j && j->lat_full ) ? 1 : 0; }
int NS(TrackJob_block_size)( const NS(TrackJobBase)* j )
{
    return j ? ( ( j->block == 64 || j->block == 128 ) ? j->block : 256 ) : 0;
}

int64_t NS(TrackJob_required_output_sets)( NS(Buffer)* pbuf, NS(buffer_size_t) nsets,
    NS(buffer_size_t) const* idx, NS(Buffer)* ebuf, NS(buffer_size_t) ute,
    int64_t* out_sizes, int64_t max_sets, int64_t* has_ebe, char const* config_str )
{
    if( pbuf == nullptr || ebuf == nullptr ) return -1;
    std::vector< uint64_t > sets;
    if( nsets == 0 || idx == nullptr ) sets.push_back( 0u ); else sets.assign( idx, idx + nsets );
    Attr a;
    OutputPlan plan;
    if( !find_attributes( pbuf, sets, ebuf, a, nullptr ) ) return -1;
    long const passes = config_int( config_str, "ebe_store_passes", 0 );
    plan_output( ebuf, a, ute, passes > 0 ? passes : 0, plan );
    std::vector< int64_t > need;
    if( plan.has_ebe ) need.push_back( plan.ebe_size );
    need.insert( need.end(), plan.monitor_sizes.begin(), plan.monitor_sizes.end() );
    if( has_ebe ) *has_ebe = plan.has_ebe ? 1 : 0;
    for( int64_t k = 0; k < ( int64_t )need.size() && k < max_sets && out_sizes; ++k )
        out_sizes[ k ] = need[ k ];
    return ( int64_t )need.size();
}

int NS(Hip_num_devices)( void )
{
    int n = 0;
    if( hipGetDeviceCount( &n ) != hipSuccess ) return 0;
    return n;
}

char const* NS(Hip_last_error_str)( void ) { return g_err.c_str(); }

// library-internal: other translation units (batch.cpp) report errors here
void st_internal_set_error( char const* msg ) { g_err = msg ? msg : ""; }

// -- stored buffers (track_job.h:717-772) --------------------------------
NS(arch_size_t) NS(TrackJob_add_stored_buffer)( NS(TrackJobBase)* j, NS(Buffer)* buffer,
    bool take_ownership, bool delete_ptr_after_move )
{
    ( void )delete_ptr_after_move;
    if( j == nullptr || buffer == nullptr ) return ST_ARCH_ILLEGAL_BUFFER_ID;
    if( NS(Buffer_needs_remapping)( buffer ) )
    { g_err = "stored buffers must not need remapping"; return ST_ARCH_ILLEGAL_BUFFER_ID; }
    j->stored.push_back( { buffer, take_ownership } );
    return ST_ARCH_MIN_USER_DEFINED_BUFFER_ID + j->stored.size() - 1u;
}

NS(arch_size_t) NS(TrackJob_create_stored_buffer)( NS(TrackJobBase)* j,
    NS(buffer_size_t) buffer_capacity )
{
    if( j == nullptr ) return ST_ARCH_ILLEGAL_BUFFER_ID;
    NS(Buffer)* b = NS(Buffer_new)( buffer_capacity );
    if( b == nullptr ) return ST_ARCH_ILLEGAL_BUFFER_ID;
    j->stored.push_back( { b, true } );
    return ST_ARCH_MIN_USER_DEFINED_BUFFER_ID + j->stored.size() - 1u;
}

NS(arch_status_t) NS(TrackJob_remove_stored_buffer)( NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    if( j == nullptr || id < ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ||
        id - ST_ARCH_MIN_USER_DEFINED_BUFFER_ID >= j->stored.size() ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    auto& b = j->stored[ id - ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ];
    if( b.buf == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    if( b.owned ) NS(Buffer_delete)( b.buf );
    b.buf = nullptr; b.owned = false;               // ids of the others stay valid
    return ST_ARCH_STATUS_SUCCESS;
}

NS(arch_size_t) NS(TrackJob_num_stored_buffers)( const NS(TrackJobBase)* j )
{
    NS(arch_size_t) n = 0;
    if( j ) for( const auto& b : j->stored ) n += b.buf != nullptr ? 1u : 0u;
    return n;
}

bool NS(TrackJob_has_stored_buffers)( const NS(TrackJobBase)* j ) { return NS(TrackJob_num_stored_buffers)( j ) > 0u; }

NS(arch_size_t) NS(TrackJob_min_stored_buffer_id)( const NS(TrackJobBase)* j )
{
    return NS(TrackJob_has_stored_buffers)( j ) ? ST_ARCH_MIN_USER_DEFINED_BUFFER_ID : ST_ARCH_ILLEGAL_BUFFER_ID;
}

NS(arch_size_t) NS(TrackJob_max_stored_buffer_id)( const NS(TrackJobBase)* j )
{
    return NS(TrackJob_has_stored_buffers)( j )
        ? ST_ARCH_MIN_USER_DEFINED_BUFFER_ID + j->stored.size() - 1u : ST_ARCH_ILLEGAL_BUFFER_ID;
}

bool NS(TrackJob_owns_stored_buffer)( const NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    if( j == nullptr || id < ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ||
        id - ST_ARCH_MIN_USER_DEFINED_BUFFER_ID >= j->stored.size() ) return false;
    auto const& b = j->stored[ id - ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ];
    return b.buf != nullptr && b.owned;
}

NS(Buffer)* NS(TrackJob_stored_buffer)( NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    return ( j && id >= ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ) ? j->buffer_by_id( id ) : nullptr;
}

// the device holds stored data only as copies inside the packed lattice:
// a push rebuilds it, a collect has nothing to bring back
NS(arch_status_t) NS(TrackJob_push_stored_buffer)( NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    if( NS(TrackJob_stored_buffer)( j, id ) == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    return ( j->d_lat == nullptr || j->pack_lattice() ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJob_collect_stored_buffer)( NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    return NS(TrackJob_stored_buffer)( j, id ) != nullptr ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

// -- address assignment items (track_job.h:500-660) -----------------------
NS(AssignAddressItem)* NS(TrackJob_add_assign_address_item)( NS(TrackJobBase)* j,
    const NS(AssignAddressItem)* item )
{
    if( j == nullptr || item == nullptr ) return nullptr;
    for( NS(AssignAddressItem)& it : j->assign_items )
        if( std::memcmp( &it, item, sizeof( it ) ) == 0 ) return &it;   // already there
    // after the last item of the same (dest, src) pair: a pair's items stay
    // next to each other, in the order added
    size_t pos = j->assign_items.size();
    for( size_t k = 0; k < j->assign_items.size(); ++k )
        if( j->assign_items[ k ].dest_buffer_id == item->dest_buffer_id &&
            j->assign_items[ k ].src_buffer_id == item->src_buffer_id ) pos = k + 1;
    j->assign_items.insert( j->assign_items.begin() + pos, *item );
    ++j->items_version;
    return &j->assign_items[ pos ];
}

NS(AssignAddressItem)* NS(TrackJob_add_assign_address_item_detailed)( NS(TrackJobBase)* j,
    NS(object_type_id_t) dest_type_id, NS(buffer_size_t) dest_buffer_id,
    NS(buffer_size_t) dest_elem_index, NS(buffer_size_t) dest_pointer_offset,
    NS(object_type_id_t) src_type_id, NS(buffer_size_t) src_buffer_id,
    NS(buffer_size_t) src_elem_index, NS(buffer_size_t) src_pointer_offset )
{
    NS(AssignAddressItem) it{};
    it.dest_buffer_id = dest_buffer_id; it.src_buffer_id = src_buffer_id;
    it.dest_elem_type_id = dest_type_id; it.src_elem_type_id = src_type_id;
    it.dest_elem_index = dest_elem_index; it.src_elem_index = src_elem_index;
    it.dest_pointer_offset = dest_pointer_offset; it.src_pointer_offset = src_pointer_offset;
    return NS(TrackJob_add_assign_address_item)( j, &it );
}

NS(arch_size_t) NS(TrackJob_num_assign_items)( const NS(TrackJobBase)* j,
    NS(arch_size_t) dest_buffer_id, NS(arch_size_t) src_buffer_id )
{
    NS(arch_size_t) n = 0;
    if( j ) for( const auto& it : j->assign_items )
        n += ( it.dest_buffer_id == dest_buffer_id && it.src_buffer_id == src_buffer_id ) ? 1u : 0u;
    return n;
}

bool NS(TrackJob_has_assign_items)( const NS(TrackJobBase)* j, NS(arch_size_t) dest_buffer_id,
    NS(arch_size_t) src_buffer_id )
{
    return NS(TrackJob_num_assign_items)( j, dest_buffer_id, src_buffer_id ) > 0u;
}

NS(arch_size_t) NS(TrackJob_total_num_assign_items)( const NS(TrackJobBase)* j )
{
    return j ? j->assign_items.size() : 0u;
}

NS(arch_status_t) NS(TrackJob_commit_address_assignments)( NS(TrackJobBase)* j )
{
    return j ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJob_assign_all_addresses)( NS(TrackJobBase)* j )
{
    return ( j && j->assign( true, 0u, 0u ) ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJob_assign_addresses)( NS(TrackJobBase)* j,
    NS(buffer_size_t) dest_buffer_id, NS(buffer_size_t) src_buffer_id )
{
    return ( j && j->assign( false, dest_buffer_id, src_buffer_id ) )
        ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}


// -- arch / device strings (track_job.h:123-135) ---------------------------
int NS(TrackJob_extract_device_id_str)( char const* conf, char* out, NS(buffer_size_t) max_len )
{
    if( conf == nullptr ) return -1;
    if( out != nullptr && max_len > 0u ) std::memset( out, 0, max_len );
    std::string id;
    std::string const c = conf;
    size_t const at = c.find( "device=" );
    if( at != std::string::npos && ( at == 0 || c[ at - 1 ] == ' ' || c[ at - 1 ] == ',' ||
                                     c[ at - 1 ] == ';' ) )
    {
        size_t e = at + 7;
        while( e < c.size() && std::isalnum( ( unsigned char )c[ e ] ) ) ++e;
        id = c.substr( at + 7, e - at - 7 );
    }
    else
    {
        // a leading "P.D" (platform.device), the reference's device id form
        size_t b = 0;
        while( b < c.size() && std::isspace( ( unsigned char )c[ b ] ) ) ++b;
        size_t e = b;
        while( e < c.size() && std::isdigit( ( unsigned char )c[ e ] ) ) ++e;
        if( e > b && e < c.size() && c[ e ] == '.' )
        {
            size_t f = e + 1;
            while( f < c.size() && std::isdigit( ( unsigned char )c[ f ] ) ) ++f;
            if( f > e + 1 ) id = c.substr( b, f - b );
        }
    }
    if( id.empty() ) return 0;
    if( out == nullptr || id.size() + 1u > max_len ) return -1;
    std::memcpy( out, id.data(), id.size() );
    return 0;
}

int NS(TrackJob_sanitize_arch_str)( char const* arch, char* out, NS(buffer_size_t) max_len )
{
    if( arch == nullptr || out == nullptr || max_len == 0u ) return -1;
    std::string a;
    for( char const* c = arch; *c; ++c )
        if( !std::isspace( ( unsigned char )*c ) ) a.push_back( ( char )std::tolower( ( unsigned char )*c ) );
    if( a.empty() || a.size() + 1u > max_len ) return -1;
    std::memset( out, 0, max_len );
    std::memcpy( out, a.data(), a.size() );
    return 0;
}

int NS(TrackJob_sanitize_arch_str_inplace)( char* arch, NS(buffer_size_t) max_len )
{
    if( arch == nullptr || std::strlen( arch ) >= max_len ) return -1;
    std::string const copy = arch;
    return NS(TrackJob_sanitize_arch_str)( copy.c_str(), arch, max_len );
}

// -- particle addresses (track_job.h:206-290) ------------------------------
void NS(TrackJob_collect_particles_addresses)( NS(TrackJobBase)* ) {}
void NS(TrackJob_push_particles_addresses)( NS(TrackJobBase)* ) {}
bool NS(TrackJob_can_fetch_particles_addr)( const NS(TrackJobBase)* j ) { return j && j->paddr_buf; }
bool NS(TrackJob_has_particles_addr)( const NS(TrackJobBase)* j ) { return j && j->has_paddr; }

NS(arch_status_t) NS(TrackJob_fetch_particle_addresses)( NS(TrackJobBase)* j )
{
    return ( j && j->fetch_paddr() ) ? ST_ARCH_STATUS_SUCCESS : ST_ARCH_STATUS_GENERAL_FAILURE;
}
NS(arch_status_t) NS(TrackJob_fetch_particles_addr)( NS(TrackJobBase)* j )
{
    return NS(TrackJob_fetch_particle_addresses)( j );
}

NS(arch_status_t) NS(TrackJob_clear_particle_addresses)( NS(TrackJobBase)* j, NS(arch_size_t) k )
{
    NS(ParticlesAddr)* a = j ? j->paddr( k ) : nullptr;
    if( a == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    std::memset( a, 0, sizeof( *a ) );
    j->has_paddr = false;
    return ST_ARCH_STATUS_SUCCESS;
}
NS(arch_status_t) NS(TrackJob_clear_particles_addr)( NS(TrackJobBase)* j, NS(buffer_size_t) k )
{
    return NS(TrackJob_clear_particle_addresses)( j, k );
}

NS(arch_status_t) NS(TrackJob_clear_all_particle_addresses)( NS(TrackJobBase)* j )
{
    if( j == nullptr || j->paddr_buf == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    for( uint64_t k = 0; k < NS(Buffer_get_num_of_objects)( j->paddr_buf ); ++k )
        std::memset( j->paddr( k ), 0, sizeof( NS(ParticlesAddr) ) );
    j->has_paddr = false;
    return ST_ARCH_STATUS_SUCCESS;
}
NS(arch_status_t) NS(TrackJob_clear_all_particles_addr)( NS(TrackJobBase)* j )
{
    return NS(TrackJob_clear_all_particle_addresses)( j );
}

NS(ParticlesAddr) const* NS(TrackJob_particle_addresses)( const NS(TrackJobBase)* j, NS(arch_size_t) k )
{
    return ( j && j->has_paddr ) ? j->paddr( k ) : nullptr;
}
NS(ParticlesAddr) const* NS(TrackJob_particles_addr)( const NS(TrackJobBase)* j, NS(buffer_size_t) k )
{
    return NS(TrackJob_particle_addresses)( j, k );
}
NS(Buffer)* NS(TrackJob_get_particles_addr_buffer)( NS(TrackJobBase)* j ) { return j ? j->paddr_buf : nullptr; }
NS(Buffer) const* NS(TrackJob_get_const_particles_addr_buffer)( const NS(TrackJobBase)* j )
{
    return j ? j->paddr_buf : nullptr;
}
NS(Buffer) const* NS(TrackJob_particles_addr_buffer)( const NS(TrackJobBase)* j ) { return j ? j->paddr_buf : nullptr; }

// -- reset of one set, set sizes, const buffers, monitors, ebe config -----
bool NS(TrackJob_reset_particle_set)( NS(TrackJobBase)* job, NS(Buffer)* pbuf, NS(buffer_size_t) k,
    NS(Buffer)* ebuf, NS(Buffer)* obuf )
{
    return NS(TrackJob_reset_detailed)( job, pbuf, 1u, &k, ebuf, obuf,
                                        job ? job->until_turn_ebe : 0u );
}

NS(buffer_size_t) const* NS(TrackJob_get_num_particles_in_sets_begin)( const NS(TrackJobBase)* j )
{
    return j ? j->set_sizes_u64.data() : nullptr;
}
NS(buffer_size_t) const* NS(TrackJob_get_num_particles_in_sets_end)( const NS(TrackJobBase)* j )
{
    return j ? j->set_sizes_u64.data() + j->set_sizes_u64.size() : nullptr;
}
NS(buffer_size_t) NS(TrackJob_get_num_particles_in_set)( const NS(TrackJobBase)* j, NS(buffer_size_t) n )
{
    return ( j && n < j->set_sizes_u64.size() ) ? j->set_sizes_u64[ n ] : 0u;
}

NS(Buffer) const* NS(TrackJob_get_const_particles_buffer)( const NS(TrackJobBase)* j ) { return j ? j->pbuf : nullptr; }
NS(Buffer) const* NS(TrackJob_get_const_beam_elements_buffer)( const NS(TrackJobBase)* j ) { return j ? j->ebuf : nullptr; }
NS(Buffer) const* NS(TrackJob_get_const_output_buffer)( const NS(TrackJobBase)* j ) { return j ? j->obuf : nullptr; }

bool NS(TrackJob_has_beam_monitors)( const NS(TrackJobBase)* j ) { return j && !j->monitor_indices_u64.empty(); }
NS(buffer_size_t) const* NS(TrackJob_get_beam_monitor_indices_begin)( const NS(TrackJobBase)* j )
{
    return j ? j->monitor_indices_u64.data() : nullptr;
}
NS(buffer_size_t) const* NS(TrackJob_get_beam_monitor_indices_end)( const NS(TrackJobBase)* j )
{
    return j ? j->monitor_indices_u64.data() + j->monitor_indices_u64.size() : nullptr;
}
NS(buffer_size_t) NS(TrackJob_get_beam_monitor_index)( const NS(TrackJobBase)* j, NS(buffer_size_t) n )
{
    return ( j && n < j->monitor_indices_u64.size() ) ? j->monitor_indices_u64[ n ] : 0u;
}

static_assert( sizeof( NS(ElemByElemConfig) ) == sizeof( st_ebe_config_t ),
               "NS(ElemByElemConfig) and the kernels' configuration share their layout" );

bool NS(TrackJob_has_elem_by_elem_config)( const NS(TrackJobBase)* j ) { return j && j->plan.has_ebe; }
NS(ElemByElemConfig) const* NS(TrackJob_get_elem_by_elem_config)( const NS(TrackJobBase)* j )
{
    return ( j && j->plan.has_ebe ) ? reinterpret_cast< NS(ElemByElemConfig) const* >( &j->ebe_cfg_host )
                                    : nullptr;
}
bool NS(TrackJob_is_elem_by_elem_config_rolling)( const NS(TrackJobBase)* j )
{
    return j && j->plan.has_ebe && j->ebe_cfg_host.is_rolling != 0;
}
bool NS(TrackJob_get_default_elem_by_elem_config_rolling_flag)( const NS(TrackJobBase)* j )
{
    return j && j->ebe_rolling;
}
void NS(TrackJob_set_default_elem_by_elem_config_rolling_flag)( NS(TrackJobBase)* j, bool flag )
{
    if( j ) j->ebe_rolling = flag;
}
NS(elem_by_elem_order_t) NS(TrackJob_get_elem_by_elem_config_order)( const NS(TrackJobBase)* j )
{
    return ( j && j->plan.has_ebe ) ? j->ebe_cfg_host.order : ST_ELEM_BY_ELEM_ORDER_DEFAULT;
}
NS(elem_by_elem_order_t) NS(TrackJob_get_default_elem_by_elem_config_order)( const NS(TrackJobBase)* )
{
    return ST_ELEM_BY_ELEM_ORDER_DEFAULT;
}
void NS(TrackJob_set_default_elem_by_elem_config_order)( NS(TrackJobBase)*, NS(elem_by_elem_order_t) order )
{
    // one order exists (turn, element, particle), as in the reference
    if( order != ST_ELEM_BY_ELEM_ORDER_DEFAULT ) g_err = "only the turn-element-particle order exists";
}

// -- stored buffers: capacity, const access; buffers by id ------------------
NS(Buffer) const* NS(TrackJob_const_stored_buffer)( const NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    return ( j && id >= ST_ARCH_MIN_USER_DEFINED_BUFFER_ID ) ? j->buffer_by_id( id ) : nullptr;
}
NS(arch_size_t) NS(TrackJob_stored_buffers_capacity)( const NS(TrackJobBase)* j )
{
    return j ? j->stored.capacity() : 0u;
}
NS(arch_status_t) NS(TrackJob_reserve_stored_buffers_capacity)( NS(TrackJobBase)* j, NS(arch_size_t) cap )
{
    if( j == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    j->stored.reserve( cap );
    return ST_ARCH_STATUS_SUCCESS;
}
NS(Buffer)* NS(TrackJob_buffer_by_buffer_id)( NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    return j ? j->buffer_by_id( id ) : nullptr;
}
NS(Buffer) const* NS(TrackJob_const_buffer_by_buffer_id)( const NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    return j ? j->buffer_by_id( id ) : nullptr;
}
bool NS(TrackJob_is_buffer_by_buffer_id)( const NS(TrackJobBase)* j, NS(arch_size_t) id )
{
    return j && j->buffer_by_id( id ) != nullptr;
}
bool NS(TrackJob_is_raw_memory_by_buffer_id)( const NS(TrackJobBase)*, NS(arch_size_t) ) { return false; }

// -- assignment items: search, removal, pairs ------------------------------
namespace
{
bool same_item( const NS(AssignAddressItem)& a, const NS(AssignAddressItem)& b )
{
    return std::memcmp( &a, &b, sizeof( a ) ) == 0;
}

NS(AssignAddressItem) make_item( NS(object_type_id_t) dt, NS(buffer_size_t) db, NS(buffer_size_t) di,
    NS(buffer_size_t) dp, NS(object_type_id_t) st, NS(buffer_size_t) sb, NS(buffer_size_t) si,
    NS(buffer_size_t) sp )
{
    NS(AssignAddressItem) it{};
    it.dest_buffer_id = db; it.src_buffer_id = sb; it.dest_elem_type_id = dt; it.src_elem_type_id = st;
    it.dest_elem_index = di; it.src_elem_index = si; it.dest_pointer_offset = dp; it.src_pointer_offset = sp;
    return it;
}

// position of item k within its (dest, src) pair and the range of the pair
void pair_range( const NS(TrackJobBase)* j, uint64_t dest, uint64_t src, size_t& b, size_t& e )
{
    b = e = 0;
    bool found = false;
    for( size_t k = 0; k < j->assign_items.size(); ++k )
    {
        bool const in = j->assign_items[ k ].dest_buffer_id == dest && j->assign_items[ k ].src_buffer_id == src;
        if( in && !found ) { b = k; found = true; }
        if( in ) e = k + 1;
    }
}
}  // namespace

NS(AssignAddressItem)* NS(TrackJob_add_assign_address_item_detail)( NS(TrackJobBase)* j,
    NS(object_type_id_t) dt, NS(arch_size_t) db, NS(arch_size_t) di, NS(arch_size_t) dp,
    NS(object_type_id_t) st, NS(arch_size_t) sb, NS(arch_size_t) si, NS(arch_size_t) sp )
{
    return NS(TrackJob_add_assign_address_item_detailed)( j, dt, db, di, dp, st, sb, si, sp );
}

NS(arch_status_t) NS(TrackJob_remove_assign_address_item)( NS(TrackJobBase)* j, NS(AssignAddressItem)* item )
{
    if( j == nullptr || item == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    for( size_t k = 0; k < j->assign_items.size(); ++k )
        if( same_item( j->assign_items[ k ], *item ) )
        {
            j->assign_items.erase( j->assign_items.begin() + k );
            ++j->items_version;
            return ST_ARCH_STATUS_SUCCESS;
        }
    return ST_ARCH_STATUS_GENERAL_FAILURE;
}

NS(arch_status_t) NS(TrackJob_remove_assign_address_item_by_key_and_index)( NS(TrackJobBase)* j,
    const NS(TrackJobDestSrcBufferIds)* key, NS(buffer_size_t) index )
{
    if( j == nullptr || key == nullptr ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    size_t b, e;
    pair_range( j, key->dest_buffer_id, key->src_buffer_id, b, e );
    if( b + index >= e ) return ST_ARCH_STATUS_GENERAL_FAILURE;
    j->assign_items.erase( j->assign_items.begin() + ( b + index ) );
    ++j->items_version;
    return ST_ARCH_STATUS_SUCCESS;
}

NS(AssignAddressItem) const* NS(TrackJob_ptr_assign_address_item)( const NS(TrackJobBase)* j,
    const NS(AssignAddressItem)* item )
{
    if( j == nullptr || item == nullptr ) return nullptr;
    for( const NS(AssignAddressItem)& it : j->assign_items ) if( same_item( it, *item ) ) return &it;
    return nullptr;
}

bool NS(TrackJob_has_assign_address_item)( const NS(TrackJobBase)* j, const NS(AssignAddressItem)* item )
{
    return NS(TrackJob_ptr_assign_address_item)( j, item ) != nullptr;
}

NS(AssignAddressItem) const* NS(TrackJob_ptr_assign_address_item_detailed)( const NS(TrackJobBase)* j,
    NS(object_type_id_t) dt, NS(buffer_size_t) db, NS(buffer_size_t) di, NS(buffer_size_t) dp,
    NS(object_type_id_t) st, NS(buffer_size_t) sb, NS(buffer_size_t) si, NS(buffer_size_t) sp )
{
    NS(AssignAddressItem) const it = make_item( dt, db, di, dp, st, sb, si, sp );
    return NS(TrackJob_ptr_assign_address_item)( j, &it );
}

bool NS(TrackJob_has_assign_address_item_detailed)( const NS(TrackJobBase)* j,
    NS(object_type_id_t) dt, NS(buffer_size_t) db, NS(buffer_size_t) di, NS(buffer_size_t) dp,
    NS(object_type_id_t) st, NS(buffer_size_t) sb, NS(buffer_size_t) si, NS(buffer_size_t) sp )
{
    return NS(TrackJob_ptr_assign_address_item_detailed)( j, dt, db, di, dp, st, sb, si, sp ) != nullptr;
}

NS(buffer_size_t) NS(TrackJob_index_of_assign_address_item)( const NS(TrackJobBase)* j,
    const NS(AssignAddressItem)* item )
{
    if( j == nullptr || item == nullptr ) return 0u;
    size_t b, e;
    pair_range( j, item->dest_buffer_id, item->src_buffer_id, b, e );
    for( size_t k = b; k < e; ++k ) if( same_item( j->assign_items[ k ], *item ) ) return k - b;
    return e - b;
}

NS(buffer_size_t) NS(TrackJob_index_of_assign_address_item_detailed)( const NS(TrackJobBase)* j,
    NS(object_type_id_t) dt, NS(buffer_size_t) db, NS(buffer_size_t) di, NS(buffer_size_t) dp,
    NS(object_type_id_t) st, NS(buffer_size_t) sb, NS(buffer_size_t) si, NS(buffer_size_t) sp )
{
    NS(AssignAddressItem) const it = make_item( dt, db, di, dp, st, sb, si, sp );
    return NS(TrackJob_index_of_assign_address_item)( j, &it );
}

NS(AssignAddressItem) const* NS(TrackJob_ptr_assign_address_item_by_index)( const NS(TrackJobBase)* j,
    NS(buffer_size_t) dest, NS(buffer_size_t) src, NS(buffer_size_t) index )
{
    if( j == nullptr ) return nullptr;
    size_t b, e;
    pair_range( j, dest, src, b, e );
    return ( b + index < e ) ? &j->assign_items[ b + index ] : nullptr;
}

bool NS(TrackJob_has_assign_item_by_index)( const NS(TrackJobBase)* j, NS(buffer_size_t) dest,
    NS(buffer_size_t) src, NS(buffer_size_t) index )
{
    return NS(TrackJob_ptr_assign_address_item_by_index)( j, dest, src, index ) != nullptr;
}

static const std::vector< NS(AssignAddressItemDestSrcPair) >& item_pairs( const NS(TrackJobBase)* j )
{
    auto& pc = const_cast< NS(TrackJobBase)* >( j )->pair_cache;
    pc.clear();
    for( const NS(AssignAddressItem)& it : j->assign_items )
    {
        bool seen = false;
        for( const auto& p : pc ) seen = seen || ( p.dest_buffer_id == it.dest_buffer_id &&
                                                   p.src_buffer_id == it.src_buffer_id );
        if( !seen ) pc.push_back( { it.dest_buffer_id, it.src_buffer_id } );
    }
    return pc;
}

NS(buffer_size_t) NS(TrackJob_num_distinct_available_assign_address_items_dest_src_pairs)(
    const NS(TrackJobBase)* j )
{
    return j ? item_pairs( j ).size() : 0u;
}
NS(arch_size_t) NS(TrackJob_num_distinct_assign_items_dest_src_pairs)( const NS(TrackJobBase)* j )
{
    return NS(TrackJob_num_distinct_available_assign_address_items_dest_src_pairs)( j );
}

NS(buffer_size_t) NS(TrackJob_available_assign_address_items_dest_src_pairs)( const NS(TrackJobBase)* j,
    NS(buffer_size_t) max_num, NS(TrackJobDestSrcBufferIds)* out )
{
    if( j == nullptr ) return 0u;
    auto const& pc = item_pairs( j );
    size_t const n = std::min< size_t >( pc.size(), max_num );
    for( size_t k = 0; k < n && out; ++k ) out[ k ] = pc[ k ];
    return n;
}
NS(arch_size_t) NS(TrackJob_assign_items_dest_src_pairs)( const NS(TrackJobBase)* j, NS(arch_size_t) max_num,
    NS(AssignAddressItemDestSrcPair)* out )
{
    return NS(TrackJob_available_assign_address_items_dest_src_pairs)( j, max_num, out );
}

NS(TrackJobDestSrcBufferIds) const* NS(TrackJob_assign_item_dest_src_begin)( const NS(TrackJobBase)* j )
{
    return j ? item_pairs( j ).data() : nullptr;
}
NS(TrackJobDestSrcBufferIds) const* NS(TrackJob_assign_item_dest_src_end)( const NS(TrackJobBase)* j )
{
    return j ? j->pair_cache.data() + j->pair_cache.size() : nullptr;
}

NS(AssignAddressItem) const* NS(TrackJob_assign_items_dest_src_begin)( const NS(TrackJobBase)* j,
    NS(arch_size_t) dest, NS(arch_size_t) src )
{
    if( j == nullptr ) return nullptr;
    size_t b, e;
    pair_range( j, dest, src, b, e );
    return j->assign_items.data() + b;
}
NS(AssignAddressItem) const* NS(TrackJob_assign_items_dest_src_end)( const NS(TrackJobBase)* j,
    NS(arch_size_t) dest, NS(arch_size_t) src )
{
    if( j == nullptr ) return nullptr;
    size_t b, e;
    pair_range( j, dest, src, b, e );
    return j->assign_items.data() + e;
}

static NS(Object) const* pair_objects( const NS(TrackJobBase)* jc, uint64_t dest, uint64_t src, bool end )
{
    if( jc == nullptr ) return nullptr;
    NS(TrackJobBase)* j = const_cast< NS(TrackJobBase)* >( jc );
    // rebuilt when the pair or the items changed, so that begin and end of
    // one pair come from the same buffer
    if( j->item_objs == nullptr || j->item_objs_version != j->items_version ||
        j->item_objs_dest != dest || j->item_objs_src != src )
    {
        if( j->item_objs ) NS(Buffer_delete)( j->item_objs );
        j->item_objs = NS(Buffer_new)( 0 );
        if( j->item_objs == nullptr ) return nullptr;
        size_t b, e;
        pair_range( j, dest, src, b, e );
        for( size_t k = b; k < e; ++k )
            if( NS(Buffer_add_object)( j->item_objs, &j->assign_items[ k ], sizeof( NS(AssignAddressItem) ),
                                       65536u, 0, nullptr, nullptr, nullptr ) == nullptr ) return nullptr;
        j->item_objs_version = j->items_version;
        j->item_objs_dest = dest;
        j->item_objs_src = src;
    }
    return end ? NS(Buffer_get_objects_end)( j->item_objs ) : NS(Buffer_get_objects_begin)( j->item_objs );
}

NS(Object) const* NS(TrackJob_assign_items_begin)( const NS(TrackJobBase)* j, NS(buffer_size_t) dest,
    NS(buffer_size_t) src )
{
    return pair_objects( j, dest, src, false );
}
NS(Object) const* NS(TrackJob_assign_items_end)( const NS(TrackJobBase)* j, NS(buffer_size_t) dest,
    NS(buffer_size_t) src )
{
    return pair_objects( j, dest, src, true );
}

}  // extern "C"

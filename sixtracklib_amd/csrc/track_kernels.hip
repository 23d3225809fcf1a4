// Tracking kernels for gfx950 (one particle per lane, state in registers).
//
// This file is compiled twice (see sixtracklib_amd/build.py):
//   ST_KSUFFIX=_exact  -ffp-contract=off : bit-identical to the CPU oracle
//   ST_KSUFFIX=_fma    -ffp-contract=fast: fused multiply-adds (performance)
//
// Kernels replace the bodies of the reference's per-launch kernels
// (track/track_kernel_impl.h:102-147 until-turn, 252-306 element by
// element, 425-474 line; CUDA launch shape cuda/kernels/track_particles.cu)
// and the per-particle loops (track/track.h:610-822).  Differences in
// structure, not in results:
//   * the lattice is the packed record stream of lattice.h read through a
//     constant-address-space pointer -> scalar loads, scalar dispatch;
//   * the 13 mutable particle fields stay in VGPRs for the whole launch;
//   * at_element is not incremented per element: within a turn it is the
//     start value plus the (wave-uniform) element counter, materialised at
//     loss, at output stores and at the end;
//   * a lost lane leaves the element loop (divergent break) and idles until
//     the wave's other lanes finish the turn, exactly as the reference stops
//     tracking it.
#include <hip/hip_runtime.h>

#include "track_device.h"
#include "track_launch.h"

#ifndef ST_KSUFFIX
#error "ST_KSUFFIX (kernel name suffix of this build variant) must be defined"
#endif
#define ST_CAT2( a, b ) a##b
#define ST_CAT( a, b ) ST_CAT2( a, b )
#define KN( n ) ST_CAT( n, ST_KSUFFIX )

using namespace st;

namespace
{

ST_DEV void load_lane( const st_dev_particles_t& P, int64_t i, Lane& L, LaneConst& C )
{
    L.x = P.x[ i ]; L.px = P.px[ i ]; L.y = P.y[ i ]; L.py = P.py[ i ];
    L.zeta = P.zeta[ i ]; L.delta = P.delta[ i ]; L.rpp = P.rpp[ i ];
    L.rvv = P.rvv[ i ]; L.psigma = P.psigma[ i ]; L.s = P.s[ i ];
    L.state = P.state[ i ];
    C.q0 = P.q0[ i ]; C.mass0 = P.mass0[ i ]; C.beta0 = P.beta0[ i ];
    C.p0c = P.p0c[ i ]; C.chi = P.chi[ i ]; C.charge_ratio = P.charge_ratio[ i ];
}

ST_DEV void store_lane( const st_dev_particles_t& P, int64_t i, const Lane& L )
{
    P.x[ i ] = L.x; P.px[ i ] = L.px; P.y[ i ] = L.y; P.py[ i ] = L.py;
    P.zeta[ i ] = L.zeta; P.delta[ i ] = L.delta; P.rpp[ i ] = L.rpp;
    P.rvv[ i ] = L.rvv; P.psigma[ i ] = L.psigma; P.s[ i ] = L.s;
    P.state[ i ] = L.state;
}

// Full 21-field record (particles.h:1695-1803 field order) into output
// set O at index k.  Read-only fields come from the input set.
ST_DEV void store_record( const st_dev_particles_t& O, int64_t k,
    const st_dev_particles_t& P, int64_t i, const Lane& L, const LaneConst& C,
    int64_t pid, int64_t at_element, int64_t at_turn )
{
    O.q0[ k ] = C.q0; O.mass0[ k ] = C.mass0; O.beta0[ k ] = C.beta0;
    O.gamma0[ k ] = P.gamma0[ i ]; O.p0c[ k ] = C.p0c; O.s[ k ] = L.s;
    O.x[ k ] = L.x; O.y[ k ] = L.y; O.px[ k ] = L.px; O.py[ k ] = L.py;
    O.zeta[ k ] = L.zeta; O.psigma[ k ] = L.psigma; O.delta[ k ] = L.delta;
    O.rpp[ k ] = L.rpp; O.rvv[ k ] = L.rvv; O.chi[ k ] = C.chi;
    O.charge_ratio[ k ] = C.charge_ratio; O.particle_id[ k ] = pid;
    O.at_element[ k ] = at_element; O.at_turn[ k ] = at_turn;
    O.state[ k ] = L.state;
}

// device copy of an output particle set header, addressed by the monitor
ST_DEV st_dev_particles_t load_set( uint64_t addr )
{
    cptr_t w = ( cptr_t )addr;
    st_dev_particles_t O;
    O.n = ( int64_t )w[ 0 ];
    double** dp = &O.q0;
    for( int k = 0; k < 17; ++k ) dp[ k ] = ( double* )w[ 1 + k ];
    int64_t** ip = &O.particle_id;
    for( int k = 0; k < 4; ++k ) ip[ k ] = ( int64_t* )w[ 18 + k ];
    return O;
}

// be_monitor/track.h:47-170
ST_DEV_NOINLINE int monitor( cptr_t e, const st_dev_particles_t& P, int64_t i,
    Lane& L, const LaneConst& C, int64_t pid, int64_t at_element, int64_t at_turn )
{
    uint64_t const out = e[ 3 ];
    if( out == 0u ) return 0;
    int64_t const num_stores = ri( e, 0 ), start = ri( e, 1 ), skip = ri( e, 2 );
    int64_t const max_pid = ri( e, 4 ), min_pid = ri( e, 5 );
    int64_t const is_rolling = ri( e, 6 ), is_turn_ordered = ri( e, 7 );
    int64_t out_idx = -1;
    int64_t const since = at_turn - start;
    if( since >= 0 && ( since % skip ) == 0 )
    {
        int64_t store_idx = since / skip;
        if( store_idx >= num_stores && is_rolling ) store_idx = store_idx % num_stores;
        if( store_idx < num_stores )
        {
            int64_t const off = pid - min_pid;
            int64_t const nstore = max_pid - min_pid + 1;
            out_idx = is_turn_ordered ? nstore * store_idx + off : off * nstore + store_idx;
        }
    }
    if( out_idx < 0 ) return 0;
    st_dev_particles_t const O = load_set( out );
    if( out_idx >= O.n ) { L.state = 0; return -1; }   // failed copy
    ( void )i;
    store_record( O, out_idx, P, i, L, C, pid, at_element, at_turn );
    return 0;
}

struct Where
{
    int64_t pid, at_element, at_turn;
};

// One element (track/track.h:209-585).  Returns the map status.
template< bool FULL >
ST_DEV int apply( uint64_t h, cptr_t e, Lane& L, const LaneConst& C,
                  const st_dev_particles_t& P, int64_t i, const Where& w )
{
    unsigned const op = ST_REC_OP( h );
    if( op == ST_OP_DRIFT ) { drift( L, rd( e, 0 ) ); limit_global( L ); return 0; }
    if( op == ST_OP_MULTIPOLE ) { multipole( L, C, e, ST_REC_AUX( h ), L.delta ); return 0; }
    switch( op )
    {
    case ST_OP_DRIFT_EXACT: drift_exact( L, rd( e, 0 ) ); limit_global( L ); return 0;
    case ST_OP_CAVITY: cavity( L, C, e ); return 0;
    case ST_OP_XYSHIFT: xyshift( L, e ); return 0;
    case ST_OP_SROTATION: srotation( L, e ); return 0;
    case ST_OP_LIMIT_RECT: limit_rect( L, e ); return 0;
    case ST_OP_LIMIT_ELLIPSE: limit_ellipse( L, e ); return 0;
    case ST_OP_LIMIT_RECT_ELLIPSE: limit_rect_ellipse( L, e ); return 0;
    case ST_OP_DIPEDGE: dipedge( L, e ); return 0;
    default: break;
    }
    if( FULL )
    {
        switch( op )
        {
        case ST_OP_MONITOR: return monitor( e, P, i, L, C, w.pid, w.at_element, w.at_turn );
        case ST_OP_BB4D: bb4d( L, C, e ); return 0;
        case ST_OP_BB6D: bb6d( L, C, e ); return 0;
        case ST_OP_SC_COASTING: sc_coasting( L, C, e ); return 0;
        case ST_OP_SC_QGAUSS: sc_qgauss( L, C, e ); return 0;
        case ST_OP_SC_INTERPOL: sc_interpol( L, C, e ); return 0;
        case ST_OP_RF_MULTIPOLE: rf_multipole( L, C, e, ST_REC_AUX( h ) ); return 0;
        case ST_OP_FAIL: return -1;
        default: break;
        }
    }
    L.state = 0;                            // unknown type: track.h:577-581
    return -1;
}

// output/elem_by_elem_config.h:663-760
ST_DEV int64_t ebe_index( const st_ebe_config_t& c, int64_t pid, int64_t at_element,
                          int64_t at_turn )
{
    int64_t out = -1;
    if( c.min_particle_id <= pid && c.max_particle_id >= pid &&
        c.min_element_id <= at_element && c.max_element_id >= at_element &&
        c.min_turn <= at_turn && ( c.max_turn >= at_turn || c.is_rolling ) )
    {
        int64_t const np = c.num_particles_to_store, ne = c.num_elements_to_store;
        int64_t const cap = np * ne * c.num_turns_to_store;
        if( c.order == 0 )
        {
            out = ( at_turn - c.min_turn ) * np * ne +
                  ( at_element - c.min_element_id ) * np + ( pid - c.min_particle_id );
            if( out >= cap && cap > 0 && c.is_rolling ) out = out % cap;
        }
    }
    return out;
}

ST_DEV void report( int32_t* status_out, int32_t st )
{
    if( st != 0 ) atomicOr( status_out, st );
}

// track/track.h:610-670 + track_kernel_impl.h:102-147
template< bool FULL >
ST_DEV void until_body( const st_dev_particles_t& P, int64_t i, cptr_t lat,
                        int64_t n_elem, int64_t until_turn, int32_t* status_out )
{
    Lane L; LaneConst C;
    load_lane( P, i, L, C );
    int64_t const start_elem = P.at_element[ i ];
    int64_t at_turn = P.at_turn[ i ];
    int64_t at_element = start_elem;
    Where w = { FULL ? P.particle_id[ i ] : 0, start_elem, at_turn };
    int32_t status = ( n_elem > 0 ) ? 0 : -1;
    bool go = ( n_elem > 0 ) && ( until_turn > at_turn ) && ( L.state == 1 );
    bool const tracked = go;
    while( go )
    {
        cptr_t r = lat;
        int64_t e = 0;
        int st = 0;
        for( ; e < n_elem; ++e )
        {
            uint64_t const h = r[ 0 ];
            if( FULL ) { w.at_element = start_elem + e; w.at_turn = at_turn; }
            st = apply< FULL >( h, r + 1, L, C, P, i, w );
            r += ST_REC_LEN( h );
            if( st != 0 || L.state != 1 ) break;
        }
        status = st;
        if( e == n_elem )
        {
            ++at_turn;
            L.s = 0.0;
            go = at_turn < until_turn;
        }
        else
        {
            at_element = start_elem + e;
            L.state = 0;
            go = false;
        }
    }
    if( tracked )
    {
        store_lane( P, i, L );
        P.at_element[ i ] = at_element;
        P.at_turn[ i ] = at_turn;
    }
    report( status_out, status );
}

// track/track.h:672-777 + track_kernel_impl.h:252-306
ST_DEV void ebe_body( const st_dev_particles_t& P, int64_t i, cptr_t lat,
                      int64_t n_elem, int64_t until_turn, const st_ebe_config_t& cfg,
                      const st_dev_particles_t& O, int32_t* status_out )
{
    Lane L; LaneConst C;
    load_lane( P, i, L, C );
    int64_t const pid = P.particle_id[ i ];
    int64_t const start_elem = P.at_element[ i ];
    int64_t at_turn = P.at_turn[ i ];
    int64_t at_element = start_elem;
    int32_t status = 0;
    bool go = ( n_elem > 0 ) && ( until_turn > at_turn ) && ( at_turn >= 0 ) &&
              ( start_elem >= 0 ) && ( pid >= 0 );
    bool const tracked = go;
    while( go )
    {
        cptr_t r = lat;
        int64_t e = 0;
        int st = 0;
        for( ; e < n_elem; ++e )
        {
            uint64_t const h = r[ 0 ];
            int64_t const k = ebe_index( cfg, pid, start_elem + e, at_turn );
            if( k >= 0 && k < O.n )
                store_record( O, k, P, i, L, C, pid, start_elem + e, at_turn );
            Where const w = { pid, start_elem + e, at_turn };
            st = apply< true >( h, r + 1, L, C, P, i, w );
            r += ST_REC_LEN( h );
            if( st != 0 || L.state != 1 ) break;
        }
        status = st;
        if( e == n_elem )
        {
            ++at_turn;
            L.s = 0.0;
            go = at_turn < until_turn;
        }
        else
        {
            at_element = start_elem + e;
            go = false;
        }
    }
    if( tracked )
    {
        store_lane( P, i, L );
        P.at_element[ i ] = at_element;
        P.at_turn[ i ] = at_turn;
    }
    report( status_out, status );
}

// track/track.h:779-822 + track_kernel_impl.h:425-474
ST_DEV void line_body( const st_dev_particles_t& P, int64_t i, cptr_t lat,
                       int64_t n_line, bool finish_turn, int32_t* status_out )
{
    Lane L; LaneConst C;
    load_lane( P, i, L, C );
    int64_t const start_elem = P.at_element[ i ];
    int64_t at_turn = P.at_turn[ i ];
    int64_t at_element = start_elem;
    Where w = { P.particle_id[ i ], start_elem, at_turn };
    int st = 0;
    bool const go = ( n_line > 0 ) && ( L.state == 1 );
    if( go )
    {
        cptr_t r = lat;
        int64_t e = 0;
        for( ; e < n_line; ++e )
        {
            uint64_t const h = r[ 0 ];
            w.at_element = start_elem + e;
            st = apply< true >( h, r + 1, L, C, P, i, w );
            r += ST_REC_LEN( h );
            if( st != 0 || L.state != 1 ) break;
        }
        at_element = start_elem + e;
    }
    if( finish_turn && st == 0 && L.state == 1 )
    {
        at_element = 0;
        ++at_turn;
        L.s = 0.0;
    }
    store_lane( P, i, L );
    P.at_element[ i ] = at_element;
    P.at_turn[ i ] = at_turn;
    report( status_out, st );
}

}  // namespace

template< bool FULL >
__global__ void __launch_bounds__( 256 ) KN( st_k_track_until )(
    st_dev_particles_t P, const uint64_t* __restrict__ lat, int64_t n_elem,
    int64_t until_turn, int32_t* status_out )
{
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    if( i < P.n ) until_body< FULL >( P, i, ( cptr_t )lat, n_elem, until_turn, status_out );
}

__global__ void __launch_bounds__( 256 ) KN( st_k_track_ebe )(
    st_dev_particles_t P, const uint64_t* __restrict__ lat, int64_t n_elem,
    int64_t until_turn, st_ebe_config_t cfg, st_dev_particles_t O,
    int32_t* status_out )
{
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    if( i < P.n ) ebe_body( P, i, ( cptr_t )lat, n_elem, until_turn, cfg, O, status_out );
}

__global__ void __launch_bounds__( 256 ) KN( st_k_track_line )(
    st_dev_particles_t P, const uint64_t* __restrict__ lat, int64_t n_line,
    int finish_turn, int32_t* status_out )
{
    int64_t const i = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    if( i < P.n ) line_body( P, i, ( cptr_t )lat, n_line, finish_turn != 0, status_out );
}

// ------------------------------------------------------------ launchers

static inline unsigned grid_for( int64_t n, int block )
{
    return ( unsigned )( ( n + block - 1 ) / block );
}

extern "C" hipError_t KN( st_launch_track_until )( const st_dev_particles_t* P,
    const uint64_t* lat, int64_t n_elem, int64_t until_turn, int full,
    int32_t* status, int block, hipStream_t stream )
{
    if( P->n <= 0 ) return hipSuccess;
    if( block <= 0 || block > 256 || ( block % 64 ) != 0 ) block = 256;
    if( full )
        hipLaunchKernelGGL( KN( st_k_track_until )< true >, dim3( grid_for( P->n, block ) ),
                            dim3( block ), 0, stream, *P, lat, n_elem, until_turn, status );
    else
        hipLaunchKernelGGL( KN( st_k_track_until )< false >, dim3( grid_for( P->n, block ) ),
                            dim3( block ), 0, stream, *P, lat, n_elem, until_turn, status );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_track_ebe )( const st_dev_particles_t* P,
    const uint64_t* lat, int64_t n_elem, int64_t until_turn,
    const st_ebe_config_t* cfg, const st_dev_particles_t* O, int32_t* status,
    int block, hipStream_t stream )
{
    if( P->n <= 0 ) return hipSuccess;
    if( block <= 0 || block > 256 || ( block % 64 ) != 0 ) block = 256;
    hipLaunchKernelGGL( KN( st_k_track_ebe ), dim3( grid_for( P->n, block ) ), dim3( block ),
                        0, stream, *P, lat, n_elem, until_turn, *cfg, *O, status );
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_track_line )( const st_dev_particles_t* P,
    const uint64_t* lat, int64_t n_line, int finish_turn, int32_t* status,
    int block, hipStream_t stream )
{
    if( P->n <= 0 ) return hipSuccess;
    if( block <= 0 || block > 256 || ( block % 64 ) != 0 ) block = 256;
    hipLaunchKernelGGL( KN( st_k_track_line ), dim3( grid_for( P->n, block ) ), dim3( block ),
                        0, stream, *P, lat, n_line, finish_turn, status );
    return hipGetLastError();
}

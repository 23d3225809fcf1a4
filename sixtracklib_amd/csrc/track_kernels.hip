// Tracking kernels for gfx950 (one particle per lane, state in registers).
//
// This file is compiled twice (see sixtracklib_amd/build.py):
//   ST_KSUFFIX=_exact  -ffp-contract=off : bit-identical to the CPU oracle
//   ST_KSUFFIX=_fma    -ffp-contract=fast -DST_FMA_BUILD: fused multiply-adds
//
// Kernels replace the bodies of the reference's per-launch kernels
// (track/track_kernel_impl.h:102-147 until-turn, 252-306 element by
// element, 425-474 line; CUDA launch shape cuda/kernels/track_particles.cu)
// and the per-particle loops (track/track.h:610-822).
//
// Execution model (differs from the reference in structure, not results):
//   * The lattice is the packed record stream of lattice.h, read through a
//     constant-address-space pointer: every lane of a wave reads the same
//     record, so the loads are scalar loads and the type dispatch is a
//     scalar branch.  The next record's header is fetched one element
//     ahead.
//   * The 11 mutable particle fields stay in VGPRs for the whole launch.
//   * The element loop is wave-uniform: a particle that is lost (or whose
//     map fails) has its state written to memory at that element, exactly
//     what the reference leaves behind, and is then masked out of every
//     later store; the wave leaves the loop when no lane is active.  No
//     lane ever branches out of the loop on its own, which keeps the
//     per-element dispatch free of divergent control flow.
//   * at_element is the lane's start value plus the wave-uniform element
//     counter and is only materialised at losses, output stores and the end.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "track_device.h"
#include "track_lean.h"
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

enum { MODE_UNTIL = 0, MODE_EBE = 1, MODE_LINE = 2 };

ST_DEV void load_lane( const st_dev_particles_t& P, int64_t i, Lane& L, LaneConst& C )
{
    L.x = P.x[ i ]; L.px = P.px[ i ]; L.y = P.y[ i ]; L.py = P.py[ i ];
    L.zeta = P.zeta[ i ]; L.delta = P.delta[ i ]; L.rpp = P.rpp[ i ];
    L.rvv = P.rvv[ i ]; L.psigma = P.psigma[ i ]; L.s = P.s[ i ];
    L.state = P.state[ i ];
    C.q0 = P.q0[ i ]; C.mass0 = P.mass0[ i ]; C.beta0 = P.beta0[ i ];
    C.p0c = P.p0c[ i ]; C.chi = P.chi[ i ]; C.charge_ratio = P.charge_ratio[ i ];
}

ST_DEV void store_lane( const st_dev_particles_t& P, int64_t i, const Lane& L,
                        int64_t at_element, int64_t at_turn )
{
    P.x[ i ] = L.x; P.px[ i ] = L.px; P.y[ i ] = L.y; P.py[ i ] = L.py;
    P.zeta[ i ] = L.zeta; P.delta[ i ] = L.delta; P.rpp[ i ] = L.rpp;
    P.rvv[ i ] = L.rvv; P.psigma[ i ] = L.psigma; P.s[ i ] = L.s;
    P.state[ i ] = L.state;
    P.at_element[ i ] = at_element;
    P.at_turn[ i ] = at_turn;
}

// store_record for record k < 2^29 of every field array: one 32-bit byte
// offset serves all 21 stores (scalar base + vector offset addressing).
// Without ``constants`` the eight fields that are fixed for a particle
// (q0, mass0, beta0, gamma0, p0c, chi, charge_ratio, particle_id) are not
// written: the caller knows the slot already holds this particle's values.
// The stores are non-temporal: the kernel never reads a record back, and
// streaming stores keep their rate with the element maps issuing between
// them (plain stores lose about 8 % of the store rate there).
ST_DEV void store_record_narrow( const st_dev_particles_t& O, int64_t k, double gamma0,
    const Lane& L, const LaneConst& C, int64_t pid, int64_t at_element, int64_t at_turn,
    bool constants )
{
    // the offset is made opaque here, so that each store forms its address
    // from the field's base (uniform) and this 32-bit offset at the store,
    // rather than the loop carrying one running 64-bit address per field
    uint32_t b = static_cast< uint32_t >( k ) * 8u;
    asm volatile( "" : "+v"( b ) );
    auto d = [ b ]( double* base, double v )
    {
        __builtin_nontemporal_store(
            v, reinterpret_cast< double* >( reinterpret_cast< char* >( base ) + b ) );
    };
    auto n = [ b ]( int64_t* base, int64_t v )
    {
        __builtin_nontemporal_store(
            v, reinterpret_cast< int64_t* >( reinterpret_cast< char* >( base ) + b ) );
    };
    if( constants )
    {
        d( O.q0, C.q0 ); d( O.mass0, C.mass0 ); d( O.beta0, C.beta0 );
        d( O.gamma0, gamma0 ); d( O.p0c, C.p0c ); d( O.chi, C.chi );
        d( O.charge_ratio, C.charge_ratio ); n( O.particle_id, pid );
    }
    d( O.s, L.s );
    d( O.x, L.x ); d( O.y, L.y ); d( O.px, L.px ); d( O.py, L.py );
    d( O.zeta, L.zeta ); d( O.psigma, L.psigma ); d( O.delta, L.delta );
    d( O.rpp, L.rpp ); d( O.rvv, L.rvv );
    n( O.at_element, at_element ); n( O.at_turn, at_turn );
    n( O.state, L.state );
}

// Elements other than Drift and Multipole (track/track.h:209-585).
// ``ok`` = false: the particle is lost at this element (state &= 0);
// ``fail`` = true: the map returned a failure status (-1).
template< bool FULL >
ST_DEV void step_other( unsigned op, unsigned aux, cptr_t e, Lane& L, const LaneConst& C,
                        bool& ok, bool& fail, bool active, const st_dev_particles_t& P,
                        int64_t i, int64_t pid, int64_t at_element, int64_t at_turn )
{
    switch( op )
    {
    case ST_OP_DRIFT_EXACT: drift_exact( L, rd( e, 0 ) ); ok = limit_global_ok( L ); return;
    case ST_OP_CAVITY: cavity( L, C, e ); return;
    case ST_OP_XYSHIFT: xyshift( L, e ); return;
    case ST_OP_SROTATION: srotation( L, e ); return;
    case ST_OP_LIMIT_RECT: ok = limit_rect_ok( L, e ); return;
    case ST_OP_LIMIT_ELLIPSE: ok = limit_ellipse_ok( L, e ); return;
    case ST_OP_LIMIT_RECT_ELLIPSE: ok = limit_rect_ellipse_ok( L, e ); return;
    case ST_OP_DIPEDGE: dipedge( L, e ); return;
    default: break;
    }
    if( FULL )
    {
        // The maps below are out of line and take the particle by
        // reference: they get copies, so that the caller's state (L, C)
        // never has its address taken and stays in registers (an escaping
        // L or C would live in scratch for the whole element loop).
        Lane t = L;
        LaneConst const ct = C;
        switch( op )
        {
        case ST_OP_MONITOR:
            if( monitor( e, P, i, t, ct, pid, at_element, at_turn, active ) )
            { ok = false; fail = true; }
            return;
        case ST_OP_BB4D: bb4d( t, ct, e ); L = t; return;
        case ST_OP_BB6D: bb6d( t, ct, e ); L = t; return;
        case ST_OP_SC_COASTING: sc_coasting( t, ct, e ); L = t; return;
        case ST_OP_SC_QGAUSS: sc_qgauss( t, ct, e ); L = t; return;
        case ST_OP_SC_INTERPOL: sc_interpol( t, ct, e ); L = t; return;
        case ST_OP_RF_MULTIPOLE: rf_multipole( t, ct, e, aux ); L = t; return;
        case ST_OP_FAIL: fail = true; return;           // disabled map
        default: break;
        }
    }
    ok = false;                                           // unknown type:
    fail = true;                                          // track.h:577-581
}

// output/elem_by_elem_config.h:663-760.  ``cap`` is the store's capacity
// in records: num_particles * num_elements * num_turns as in the reference,
// or a smaller multiple of num_particles for a store that keeps only the
// most recent element passes of every particle (rolling).
ST_DEV int64_t ebe_index( const st_ebe_config_t& c, int64_t cap, int64_t pid,
                          int64_t at_element, int64_t at_turn )
{
    int64_t out = -1;
    if( c.min_particle_id <= pid && c.max_particle_id >= pid &&
        c.min_element_id <= at_element && c.max_element_id >= at_element &&
        c.min_turn <= at_turn && ( c.max_turn >= at_turn || c.is_rolling ) )
    {
        int64_t const np = c.num_particles_to_store, ne = c.num_elements_to_store;
        if( c.order == 0 )
        {
            out = ( at_turn - c.min_turn ) * np * ne +
                  ( at_element - c.min_element_id ) * np + ( pid - c.min_particle_id );
            if( out >= cap && cap > 0 && c.is_rolling ) out = out % cap;
        }
    }
    return out;
}

struct Args
{
    cptr_t lat;
    int64_t n_elem;          // elements per turn (until / ebe) or in the line
    int64_t until_turn;
    bool finish_turn;        // line mode
    st_ebe_config_t cfg;     // ebe mode
    int64_t ebe_cap;         // ebe store capacity in records
    int64_t ebe_slots;       // store slots per particle (capacity / particles)
    int64_t turn_limit;      // ebe: turns a lane may advance in this launch (0: to until_turn)
    int64_t win_base;        // ebe: lane t tracks particle (win_base + t) mod n, t < win_n
    int64_t win_n;           //      (win_n = 0: particle t)
    int32_t ebe_lanes;       // ebe: 0 every particle, 1 those not alive (state != 1), 2 alive
    st_dev_particles_t O;    // ebe output set
    int32_t* status_out;
};

// Per-particle loops: track/track.h:610-670 (until), 672-777 (element by
// element), 779-822 (line), with the kernel bodies of
// track_kernel_impl.h:102-147, 252-306, 425-474.
template< int MODE, bool FULL >
ST_DEV void track_body( const st_dev_particles_t& P, int64_t i, const Args& a )
{
    bool const in_range = i < P.n;
    int64_t const li = in_range ? i : 0;          // out-of-range lanes never store
    Lane L; LaneConst C;
    load_lane( P, li, L, C );
    int64_t const start = P.at_element[ li ];
    int64_t at_turn = P.at_turn[ li ];
    int64_t const pid = ( FULL || MODE == MODE_EBE ) ? P.particle_id[ li ] : 0;
    int const n_elem = ( int )a.n_elem;           // host guarantees < 2^31

    bool active0;
    int32_t status = 0;
    // a launch may cover only part of the requested turns (element by
    // element in windows); the lane then stops at a turn boundary exactly as
    // if until_turn were turn_stop
    int64_t const turn_stop = ( MODE == MODE_EBE && a.turn_limit > 0 &&
                                a.until_turn - at_turn > a.turn_limit ) ? at_turn + a.turn_limit
                                                                        : a.until_turn;
    if( MODE == MODE_EBE )
    {
        active0 = in_range && n_elem > 0 && a.until_turn > at_turn && at_turn >= 0 &&
                  start >= 0 && pid >= 0;
        if( a.ebe_lanes == 1 ) active0 = active0 && L.state != 1;
        else if( a.ebe_lanes == 2 ) active0 = active0 && L.state == 1;
    }
    else if( MODE == MODE_LINE )
        active0 = in_range && n_elem > 0 && L.state == 1;
    else
    {
        active0 = in_range && n_elem > 0 && a.until_turn > at_turn && L.state == 1;
        if( n_elem <= 0 ) status = -1;            // empty lattice: failure
    }

    // element by element: the store slot advances by num_particles per
    // element pass (wrapping at the capacity when rolling), so the full
    // index is computed once per turn, at the first pass inside the
    // configured element range; records below 2^29 use 32-bit offsets
    double const gamma0 = MODE == MODE_EBE ? P.gamma0[ li ] : 0.0;
    bool const narrow = MODE == MODE_EBE && a.O.n <= ( int64_t( 1 ) << 29 );
    bool const pid_ok = MODE == MODE_EBE && a.cfg.min_particle_id <= pid &&
                        a.cfg.max_particle_id >= pid;
    // A rolling store whose configuration covers every pass of this lane
    // (its particle id, all its elements, its turns from the first on)
    // gives the lane one fixed column of slots, visited in order: after
    // the first ebe_slots passes of the call every slot holds this
    // particle's fixed fields, and later records skip them.
    bool const every_pass = pid_ok && a.cfg.is_rolling && a.cfg.min_turn <= at_turn &&
                            a.cfg.min_element_id <= start &&
                            a.cfg.max_element_id >= start + n_elem - 1;
    int64_t pass = 0;                             // element passes of this call (wave)

    // lanes still being tracked; a lane leaves the mask exactly once, after
    // its final state has been stored
    uint64_t act = ballot( active0 );
    while( act != 0u )
    {
        cptr_t r = a.lat;
        uint64_t h = r[ 0 ];
        // element by element (one wave per SIMD, nothing to hide a scalar
        // load's latency behind): a record's first parameter word, a
        // drift's length, is read with its header, one element ahead
        uint64_t p0 = MODE == MODE_EBE ? r[ ST_REC_PAYLOAD ] : 0u;
        int e_end = n_elem;
        bool const turn_ok = pid_ok && a.cfg.min_turn <= at_turn &&
                             ( a.cfg.max_turn >= at_turn || a.cfg.is_rolling );
        int64_t slot = -1;                        // store slot of the next pass
        for( int e = 0; e < e_end; ++e )
        {
            cptr_t const nr = r + ST_REC_LEN( h );
            uint64_t const hn = nr[ 0 ];          // next header (stream is padded)
            uint64_t const pn = MODE == MODE_EBE ? nr[ ST_REC_PAYLOAD ] : 0u;
            unsigned const op = ST_REC_OP( h );
            cptr_t const prm = r + ST_REC_PAYLOAD;
            if( MODE == MODE_EBE )
            {
                int64_t const el = start + e;
                int64_t k = -1;
                if( turn_ok && a.cfg.min_element_id <= el && a.cfg.max_element_id >= el )
                {
                    if( slot < 0 ) slot = ebe_index( a.cfg, a.ebe_cap, pid, el, at_turn );
                    k = slot;
                    if( slot >= 0 )
                    {
                        slot += a.cfg.num_particles_to_store;
                        if( a.cfg.is_rolling && slot >= a.ebe_cap ) slot -= a.ebe_cap;
                    }
                }
                if( lane_in( act ) && k >= 0 && k < a.O.n )
                {
                    if( narrow )
                        store_record_narrow( a.O, k, gamma0, L, C, pid, el, at_turn,
                                             !every_pass || pass < a.ebe_slots );
                    else store_record( a.O, k, P, i, L, C, pid, el, at_turn );
                }
                ++pass;
            }
            uint64_t bad = 0u, failm = 0u;        // lost here / map failed
            if( op == ST_OP_DRIFT )
            {
                drift( L, MODE == MODE_EBE ? __longlong_as_double( ( long long )p0 ) : rd( prm, 0 ) );
                bad = ballot( !limit_global_ok( L ) );
            }
            else if( op == ST_OP_MULTIPOLE )
            {
                multipole( L, C, prm, ST_REC_AUX( h ), L.delta );
            }
            else
            {
                bool ok = true, fail = false;
                step_other< FULL >( op, ST_REC_AUX( h ), prm, L, C, ok, fail,
                                    lane_in( act ), P, i, pid, start + e, at_turn );
                bad = ballot( !ok );
                failm = ballot( fail );
            }
            if( MODE == MODE_EBE ) bad |= ballot( L.state != 1 );
            uint64_t const stop = ( bad | failm ) & act;
            if( stop != 0u )
            {
                if( lane_in( stop ) )
                {
                    if( lane_in( bad ) || MODE == MODE_UNTIL ) L.state = 0;
                    store_lane( P, i, L, start + e, at_turn );
                    status = lane_in( failm ) ? -1 : 0;
                }
                act &= ~stop;
                if( act == 0u ) e_end = 0;
            }
            r = nr;
            h = hn;
            p0 = pn;
        }
        if( MODE == MODE_LINE ) break;
        if( lane_in( act ) ) { ++at_turn; L.s = 0.0; }
        uint64_t const fin = ballot( at_turn >= turn_stop ) & act;
        if( fin != 0u )
        {
            if( lane_in( fin ) ) store_lane( P, i, L, start, at_turn );
            act &= ~fin;
        }
    }
    if( MODE == MODE_LINE && lane_in( act ) )
    {
        int64_t at_element = start + n_elem;
        if( a.finish_turn ) { at_element = 0; ++at_turn; L.s = 0.0; }
        store_lane( P, i, L, at_element, at_turn );
    }
    if( in_range && status != 0 ) atomicOr( a.status_out, status );
}

ST_DEV void track_until_lean( const st_dev_particles_t& P, int64_t tid, const Args& a )
{
    track_lean_lane( PDirect{ P }, nullptr, P.n, tid, a.lat, a.n_elem, a.until_turn, 0, a.status_out );
}

}  // namespace

__global__ void __launch_bounds__( 256 ) KN( st_k_track_until )( st_dev_particles_t P, Args a )
{
    track_body< MODE_UNTIL, true >( P, ( int64_t )blockIdx.x * blockDim.x + threadIdx.x, a );
}

// Occupancy target of the lean kernel (waves per SIMD): the register
// budget it implies spills only on the Cavity and store paths.
#ifndef ST_LEAN_WAVES_PER_EU
#define ST_LEAN_WAVES_PER_EU 8
#endif

__global__ void __launch_bounds__( 256, ST_LEAN_WAVES_PER_EU )
KN( st_k_track_until_lean )( st_dev_particles_t P, Args a )
{
    track_until_lean( P, ( int64_t )blockIdx.x * blockDim.x + threadIdx.x, a );
}

// One wave per workgroup and, in the launcher's windows, one per SIMD: the
// whole register file is this wave's, so the full map set keeps its state
// and the record fields in registers (no scratch spills on the store path).
__global__ void __launch_bounds__( 64 ) __attribute__( ( amdgpu_waves_per_eu( 1, 1 ) ) )
KN( st_k_track_ebe )( st_dev_particles_t P, Args a )
{
    int64_t const t = ( int64_t )blockIdx.x * blockDim.x + threadIdx.x;
    int64_t const i = a.win_n <= 0 ? t : ( t < a.win_n ? ( a.win_base + t ) % P.n : P.n );
    track_body< MODE_EBE, true >( P, i, a );
}

__global__ void __launch_bounds__( 256 ) KN( st_k_track_line )( st_dev_particles_t P, Args a )
{
    track_body< MODE_LINE, true >( P, ( int64_t )blockIdx.x * blockDim.x + threadIdx.x, a );
}

// ------------------------------------------------------------ launchers

static inline unsigned grid_for( int64_t n, int block )
{
    return ( unsigned )( ( n + block - 1 ) / block );
}

static inline int sane_block( int block )
{
    return ( block <= 0 || block > 256 || ( block % 64 ) != 0 ) ? 256 : block;
}

extern "C" hipError_t KN( st_launch_track_until )( const st_dev_particles_t* P,
    const uint64_t* lat, int64_t n_elem, int64_t until_turn, int full,
    int32_t* status, int block, hipStream_t stream )
{
    if( P->n <= 0 ) return hipSuccess;
    block = sane_block( block );
    Args a{};
    a.lat = ( cptr_t )lat; a.n_elem = n_elem; a.until_turn = until_turn;
    a.status_out = status;
    if( full )
    {
        hipLaunchKernelGGL( KN( st_k_track_until ), dim3( grid_for( P->n, block ) ),
                            dim3( block ), 0, stream, *P, a );
    }
    else
    {
        hipLaunchKernelGGL( KN( st_k_track_until_lean ), dim3( grid_for( P->n, block ) ),
                            dim3( block ), 0, stream, *P, a );
    }
    return hipGetLastError();
}

extern "C" hipError_t KN( st_launch_track_ebe )( const st_dev_particles_t* P,
    const uint64_t* lat, int64_t n_elem, int64_t until_turn,
    const st_ebe_config_t* cfg, int64_t cap, const st_dev_particles_t* O, int32_t* status,
    int64_t from_turn, int block, hipStream_t stream )
{
    if( P->n <= 0 ) return hipSuccess;
    ( void )block;
    Args a{};
    a.lat = ( cptr_t )lat; a.n_elem = n_elem; a.until_turn = until_turn;
    a.cfg = *cfg; a.ebe_cap = cap; a.O = *O; a.status_out = status;
    a.ebe_slots = cfg->num_particles_to_store > 0 ? cap / cfg->num_particles_to_store : 0;
    // The record stores bound this kernel, and a SIMD holding two of its
    // waves takes twice as long as one holding one: the store rate of a
    // SIMD does not grow with its waves.  So every launch holds at most one
    // wave per SIMD (W = 64 particles x 4 SIMDs x CUs), one wave per
    // workgroup so that the waves spread over all CUs:
    //   1. the particles not alive on entry, in windows of W: each stores
    //      its record at its element and stops there, as in one launch;
    //   2. with from_turn (the smallest at_turn of the live particles) known,
    //      L = ceil(n (until - from) / W) launches of one turn each over
    //      windows of w <= W particles (n (until - from) / L rounded up to
    //      whole waves) that advance by w modulo the set: every particle gets
    //      at least its turns (L w >= n (until - from); a particle at
    //      until_turn is left alone), and the launches are evenly filled,
    //      where rounds of the whole request would end on a part-filled one;
    //   without from_turn, the live particles in rounds of W for the whole
    //   request.
    // Every particle is tracked on its own, turn by turn, and its records'
    // slots follow from its id, element and turn, so these launches give
    // the results of one launch.  A particle lost in one launch is not alive
    // in the next.
    int n_cu = 0, dev = 0;
    if( hipGetDevice( &dev ) != hipSuccess ||
        hipDeviceGetAttribute( &n_cu, hipDeviceAttributeMultiprocessorCount, dev ) != hipSuccess ||
        n_cu <= 0 ) n_cu = 1;
    int64_t const W = 64 * 4 * ( int64_t )n_cu;
    auto launch = [ & ]( int64_t base, int64_t n, int lanes, int64_t turn_limit ) -> hipError_t
    {
        a.win_base = base; a.win_n = n; a.ebe_lanes = lanes; a.turn_limit = turn_limit;
        hipLaunchKernelGGL( KN( st_k_track_ebe ), dim3( grid_for( n, 64 ) ), dim3( 64 ), 0, stream,
                            *P, a );
        return hipGetLastError();
    };
    for( int64_t begin = 0; begin < P->n; begin += W )
    {
        hipError_t const e = launch( begin, std::min( W, P->n - begin ), 1, 0 );
        if( e != hipSuccess ) return e;
    }
    if( from_turn < 0 || from_turn >= until_turn )
    {
        for( int64_t begin = 0; from_turn < 0 && begin < P->n; begin += W )
        {
            hipError_t const e = launch( begin, std::min( W, P->n - begin ), 2, 0 );
            if( e != hipSuccess ) return e;
        }
        return hipSuccess;
    }
    // the fewest launches of at most W particles, then the window shrunk to
    // spread the particle-turns evenly over them (whole waves), so that the
    // last launch is not a few waves at the latency floor
    int64_t const turns = until_turn - from_turn;
    int64_t const launches = ( P->n * turns + std::min( W, P->n ) - 1 ) / std::min( W, P->n );
    int64_t const even = ( P->n * turns + launches - 1 ) / launches;
    int64_t const win = std::min( std::min( W, P->n ), 64 * ( ( even + 63 ) / 64 ) );
    for( int64_t j = 0; j < launches; ++j )
    {
        hipError_t const e = launch( ( j * win ) % P->n, win, 2, 1 );
        if( e != hipSuccess ) return e;
    }
    return hipSuccess;
}

extern "C" hipError_t KN( st_launch_track_line )( const st_dev_particles_t* P,
    const uint64_t* lat, int64_t n_line, int finish_turn, int32_t* status,
    int block, hipStream_t stream )
{
    if( P->n <= 0 ) return hipSuccess;
    block = sane_block( block );
    Args a{};
    a.lat = ( cptr_t )lat; a.n_elem = n_line; a.finish_turn = finish_turn != 0;
    a.status_out = status;
    hipLaunchKernelGGL( KN( st_k_track_line ), dim3( grid_for( P->n, block ) ), dim3( block ),
                        0, stream, *P, a );
    return hipGetLastError();
}

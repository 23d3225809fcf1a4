// Lean until-turn tracking of one lane (device code), shared by the plain
// kernels (track_kernels.hip) and the batched multi-seed kernel
// (batch_kernels.hip).
//
// Lattices made only of Drift, DriftExact, Multipole, Cavity, XYShift,
// SRotation, the three apertures and DipoleEdge (ST_OP_IS_LEAN) run through
// this loop.  Results are those of track/track.h:610-670 (the reference's
// per-particle until-turn loop); the structure is what differs:
//   * the element loop is two nested loops: an inner one over runs of Drift
//     and Multipole records (96 % of a typical ring), and one visit of the
//     outer loop per other record.  Only the inner loop's two maps merge
//     into the loop-carried lane state, so that state stays in place in the
//     same registers instead of being copied at every element;
//   * the inner loop tests one header bit to stay on the hot records and a
//     second to tell Drift from Multipole; records are addressed by a
//     32-bit byte offset (one scalar add per element) and the stream ends
//     with an END record, so there is no element counter;
//   * only chi is held in a register; the other per-particle constants are
//     read from memory by the (rare) Cavity;
//   * the particle state is not carried: an active lane is alive (state 1)
//     and a stored loss writes state 0.
#pragma once

#include "track_device.h"
#include "track_heavy.h"

namespace st
{

// Store of a finished lane.  The index goes through an empty asm statement
// so that the field addresses are formed on this (rare) path instead of
// being kept live in registers across the element loop.
ST_DEV void store_lane_state( const st_dev_particles_t& P, int64_t i, const Lane& L,
                              int64_t state, int64_t at_element, int64_t at_turn )
{
    int64_t j = i;
    asm volatile( "" : "+v"( j ) );
    P.x[ j ] = L.x; P.px[ j ] = L.px; P.y[ j ] = L.y; P.py[ j ] = L.py;
    P.zeta[ j ] = L.zeta; P.delta[ j ] = L.delta; P.rpp[ j ] = L.rpp;
    P.rvv[ j ] = L.rvv; P.psigma[ j ] = L.psigma; P.s[ j ] = L.s;
    P.state[ j ] = state;
    P.at_element[ j ] = at_element;
    P.at_turn[ j ] = at_turn;
}

// Access to the particle set of a lean loop.  PDirect holds the view
// (kernel arguments); PDesc reads it from a descriptor in constant memory
// at each use (count and 21 field addresses, the st_dev_particles_t
// layout), so that none of the field addresses stays live across the
// element loop: the heavy loop's out-of-line calls would otherwise force
// them to be saved and restored around every element.
struct PDirect
{
    const st_dev_particles_t& P;
    ST_DEV const st_dev_particles_t& view() const { return P; }
};

struct PDesc
{
    cptr_t d;
    ST_DEV st_dev_particles_t view() const { return load_set( ( uint64_t )d ); }
};

// Records other than Drift / Multipole.  Returns the survival predicate.
// Without HEAVY an op outside the lean set is a map failure (track.h:
// 577-581: lost, status -1); with HEAVY the other maps run out of line
// (heavy_step) on a copy of the lane, so the loop's own state stays in
// registers.
template< bool HEAVY, class PA >
ST_DEV bool lean_other( unsigned op, unsigned aux, cptr_t e, Lane& L,
                        const PA& pa, int64_t li, bool active,
                        int64_t at_element, int64_t at_turn, bool& fail )
{
    switch( op )
    {
    case ST_OP_DRIFT_EXACT: drift_exact( L, rd( e, 0 ) ); return limit_global_ok( L );
    case ST_OP_CAVITY:
    {
        int64_t j = li;
        asm volatile( "" : "+v"( j ) );
        auto const& P = pa.view();
        LaneConst C;
        C.q0 = P.q0[ j ]; C.mass0 = P.mass0[ j ]; C.beta0 = P.beta0[ j ];
        C.p0c = P.p0c[ j ]; C.chi = P.chi[ j ]; C.charge_ratio = P.charge_ratio[ j ];
        cavity( L, C, e );
        return true;
    }
    case ST_OP_XYSHIFT: xyshift( L, e ); return true;
    case ST_OP_SROTATION: srotation( L, e ); return true;
    case ST_OP_LIMIT_RECT: return limit_rect_ok( L, e );
    case ST_OP_LIMIT_ELLIPSE: return limit_ellipse_ok( L, e );
    case ST_OP_LIMIT_RECT_ELLIPSE: return limit_rect_ellipse_ok( L, e );
    case ST_OP_DIPEDGE: dipedge( L, e ); return true;
    default: break;
    }
    if constexpr( HEAVY )
    {
        Lane io = L;
        bool const ok = heavy_step( op, aux, e, &io, pa.d, li, active, at_element, at_turn, &fail );
        L = io;
        return ok;
    }
    fail = true;
    return false;
}

// Record at byte offset ``off`` of the stream (32-bit offsets: one scalar
// add per element and the offset form of the scalar load).
ST_DEV cptr_t rec_at( cptr_t lat, uint32_t off )
{
    return ( cptr_t )( ( const __attribute__( ( address_space( 4 ) ) ) char* )lat + off );
}

// Issue priority of a wave in a persistent launch.  Among waves of equal
// priority a SIMD issues the oldest first, so with equal priorities the
// waves of a SIMD progress unequally, and the last of them finish with few
// waves beside them (latency-bound).  Two rules, both wave-uniform and
// applied at turn ends:
//   progress == nullptr: by the turns the wave has left of the k the launch
//     gives a tile, in four bands, more left = higher (the waves that will
//     finish last, a tile started late or a slow one, go first);
//   progress != nullptr: every wave adds its turns to *progress, and a wave
//     behind the mean of the n_waves waves of the launch goes before one
//     ahead of it (the waves stay within about a turn of each other).
// k <= 0 leaves the priority alone.
struct WavePrio
{
    unsigned long long* progress = nullptr;
    int64_t n_waves = 0;
    int64_t k = 0;
};

ST_DEV void set_wave_priority( int p )
{
    if( p >= 3 ) __builtin_amdgcn_s_setprio( 3 );
    else if( p == 2 ) __builtin_amdgcn_s_setprio( 2 );
    else if( p == 1 ) __builtin_amdgcn_s_setprio( 1 );
    else __builtin_amdgcn_s_setprio( 0 );
}

// at the start of a tile (done = 0) and after each of its turns
ST_DEV void wave_priority( const WavePrio& wp, int64_t done )
{
    if( wp.k <= 0 ) return;
    if( wp.progress == nullptr )
    {
        set_wave_priority( ( int )( ( 4 * ( wp.k - done ) - 1 ) / wp.k ) );
        return;
    }
    if( done == 0 ) { set_wave_priority( 2 ); return; }
    unsigned long long v = 0;
    if( __lane_id() == 0 ) v = atomicAdd( wp.progress, 1ull ) + 1ull;
    v = __shfl( v, 0 );
    int64_t const g = ( int64_t )( ( uint64_t )( uint32_t )__builtin_amdgcn_readfirstlane( ( int )( uint32_t )v ) |
                                   ( ( uint64_t )( uint32_t )__builtin_amdgcn_readfirstlane( ( int )( v >> 32 ) ) << 32 ) );
    int64_t const d = g - wp.n_waves * done;                 // > 0: behind the mean
    set_wave_priority( d > wp.n_waves ? 3 : ( d > 0 ? 2 : ( d > -wp.n_waves ? 1 : 0 ) ) );
}

// a wave leaving its tile after done of k turns (its particles lost or at
// their last turn) counts the rest as done, so that the mean stays that of
// the waves still tracking
ST_DEV void wave_priority_leave( const WavePrio& wp, int64_t done )
{
    if( wp.k <= 0 || wp.progress == nullptr || done >= wp.k ) return;
    if( __lane_id() == 0 ) atomicAdd( wp.progress, ( unsigned long long )( wp.k - done ) );
}

// One turn of a lean until-turn lane through the record stream: the
// element loop of track_lean_lane.  LANE_S: the lane carries its path
// length s (added at every drift); without it every active lane started the
// turn at s = +0 and s after a drift is the record's (lattice.h), written to
// the lane only when it is lost.  The next record's header is read one
// record ahead; a record's payload is read when the record runs (reading a
// drift's two payload words ahead as well measured 4 % slower on a 20,000-
// element ring: the loop then waits at every record for two loads).
template< bool HEAVY, bool LANE_S, class PA >
ST_DEV void lean_turn( const PA& pa, int64_t i, int64_t li, cptr_t lat, Lane& L, double chi,
                       uint64_t& act, int32_t& status, int64_t start, int64_t at_turn )
{
    uint64_t s_rec = 0;                       // s after the last drift (bits)
    uint32_t off = 0;
    uint64_t h = rec_at( lat, 0 )[ 0 ];
    for( ;; )
    {
        while( ( ( uint32_t )h & ST_OP_RARE_BIT ) == 0u )   // Drift / Multipole run
        {
            cptr_t const r = rec_at( lat, off );
            uint32_t const noff = off + ST_REC_BYTES( h );
            uint64_t const hn = rec_at( lat, noff )[ 0 ];
            if( ( uint32_t )h & ST_OP_MP_BIT )
            {
                multipole( L, chi, r + ST_REC_PAYLOAD, ST_REC_AUX( h ), L.delta );
            }
            else
            {
                double const length = rd( r, 2 );
                drift_xz( L, length );
                if constexpr( LANE_S ) L.s += length;
                else s_rec = r[ 3 ];
                uint64_t const lost = limit_global_lost_mask( L ) & act;
                if( lost != 0u )
                {
                    if( lane_in( lost ) )
                    {
                        if constexpr( !LANE_S ) L.s = __longlong_as_double( ( long long )s_rec );
                        store_lane_state( pa.view(), i, L, 0, start + ri( r, 1 ), at_turn );
                    }
                    act &= ~lost;
                }
            }
            off = noff;
            h = hn;
        }
        unsigned const op = ST_REC_OP( h );
        if( op == ST_OP_END ) break;
        cptr_t const r = rec_at( lat, off );
        uint32_t const noff = off + ST_REC_BYTES( h );
        uint64_t const hn = rec_at( lat, noff )[ 0 ];
        bool fail = false;
        bool const ok = lean_other< HEAVY >( op, ST_REC_AUX( h ), r + ST_REC_PAYLOAD, L, pa, li,
                                             lane_in( act ), start + ri( r, 1 ), at_turn, fail );
        if constexpr( !LANE_S )
        {
            if( op == ST_OP_DRIFT_EXACT )
            {
                L.s = 0.0;                    // drift_exact added the length
                s_rec = r[ 3 ];
            }
        }
        uint64_t const lost = ballot( !ok ) & act;
        if( lost != 0u )
        {
            if( lane_in( lost ) )
            {
                if constexpr( !LANE_S ) L.s = __longlong_as_double( ( long long )s_rec );
                store_lane_state( pa.view(), i, L, 0, start + ri( r, 1 ), at_turn );
                if( fail ) status = -1;
            }
            act &= ~lost;
        }
        off = noff;
        h = hn;
    }
}

// One lane of a lean until-turn launch (HEAVY: the lattice may also hold
// the other maps, see lean_other).
//   P          particle set; particle i = list[tid] with a list, else tid
//   n_lanes    number of valid tid (list length, or P.n without a list)
//   lat        packed lean lattice of n_elem records (END-terminated)
//   max_turns  turns this launch may advance a lane (<= 0: until_turn)
//   wp         issue priority of the wave (wave_priority)
// The OR of the per-particle statuses goes to *status_out.
template< bool HEAVY = false, class PA = PDirect >
ST_DEV void track_lean_lane( const PA& pa, const int32_t* list, int64_t n_lanes,
                             int64_t tid, cptr_t lat, int64_t n_elem, int64_t until_turn,
                             int64_t max_turns, int32_t* status_out, const WavePrio& wp = WavePrio() )
{
    bool const in_range = tid < n_lanes;
    int64_t const i = ( list != nullptr && in_range ) ? ( int64_t )list[ tid ] : tid;
    int64_t const li = in_range ? i : 0;          // out-of-range lanes never store
    Lane L;
    double chi;
    int64_t start, at_turn;
    bool active0;
    {
        auto const& P = pa.view();
        L.x = P.x[ li ]; L.px = P.px[ li ]; L.y = P.y[ li ]; L.py = P.py[ li ];
        L.zeta = P.zeta[ li ]; L.delta = P.delta[ li ]; L.rpp = P.rpp[ li ];
        L.rvv = P.rvv[ li ]; L.psigma = P.psigma[ li ]; L.s = P.s[ li ];
        chi = P.chi[ li ];
        start = P.at_element[ li ];
        at_turn = P.at_turn[ li ];
        active0 = in_range && n_elem > 0 && until_turn > at_turn && P.state[ li ] == 1;
    }
    int32_t status = ( in_range && n_elem <= 0 ) ? -1 : 0;
    // a launch may cover only part of the requested turns; the lane then
    // stops at a turn boundary exactly as if until_turn were turn_stop
    int64_t const turn_stop = ( max_turns > 0 && until_turn - at_turn > max_turns )
                                  ? at_turn + max_turns : until_turn;

    // A lane leaves ``act`` once, after its final state has been stored.  A
    // wave whose lanes are all lost finishes the turn without effect.
    uint64_t act = ballot( active0 );
    int64_t turns_done = 0;                   // by this wave, in this call
    wave_priority( wp, 0 );
    while( act != 0u )
    {
        // A turn on which every lane starts at s = +0 (every turn after a
        // call's first, and the first when the particles start their turn
        // there) needs no per-lane path length: after each drift it is the
        // record's s (lattice.h), kept in a scalar register and written only
        // when a lane is lost.  Lattices with monitors (HEAVY) store whole
        // particles mid-turn and keep the per-lane s.
        if constexpr( HEAVY )
            lean_turn< true, true >( pa, i, li, lat, L, chi, act, status, start, at_turn );
        else if( ballot( lane_in( act ) && __double_as_longlong( L.s ) != 0 ) != 0u )
            lean_turn< false, true >( pa, i, li, lat, L, chi, act, status, start, at_turn );
        else
            lean_turn< false, false >( pa, i, li, lat, L, chi, act, status, start, at_turn );
        if( lane_in( act ) ) { ++at_turn; L.s = 0.0; }
        uint64_t const fin = ballot( at_turn >= turn_stop ) & act;
        if( fin != 0u )
        {
            if( lane_in( fin ) ) store_lane_state( pa.view(), i, L, 1, start, at_turn );
            act &= ~fin;
        }
        ++turns_done;
        if( act != 0u ) wave_priority( wp, turns_done );
    }
    wave_priority_leave( wp, turns_done );
    if( in_range && status != 0 ) atomicOr( status_out, status );
}

// Store of a finished lane of the hot-state loop: delta and psigma are not
// held in registers there (memory holds their current values), s is the
// given one.
ST_DEV void store_hot_state( const st_dev_particles_t& P, int64_t i, const Lane& L, double s,
                             int64_t state, int64_t at_element, int64_t at_turn )
{
    int64_t j = i;
    asm volatile( "" : "+v"( j ) );
    P.x[ j ] = L.x; P.px[ j ] = L.px; P.y[ j ] = L.y; P.py[ j ] = L.py;
    P.zeta[ j ] = L.zeta; P.rpp[ j ] = L.rpp; P.rvv[ j ] = L.rvv; P.s[ j ] = s;
    P.state[ j ] = state;
    P.at_element[ j ] = at_element;
    P.at_turn[ j ] = at_turn;
}

// Records other than Drift / Multipole for the hot-state loop: the maps
// that read delta (Cavity, DriftExact) take it from memory, and Cavity
// writes delta and psigma back.
template< class PA >
ST_DEV bool lean_other_hot( unsigned op, cptr_t e, Lane& L, const PA& pa, int64_t i,
                            bool active, bool& fail )
{
    switch( op )
    {
    case ST_OP_DRIFT_EXACT:
    {
        int64_t j = i;
        asm volatile( "" : "+v"( j ) );
        auto const& P = pa.view();
        L.delta = P.delta[ j ];
        L.s = 0.0;
        drift_exact( L, rd( e, 0 ) );
        return limit_global_ok( L );
    }
    case ST_OP_CAVITY:
    {
        int64_t j = i;
        asm volatile( "" : "+v"( j ) );
        auto const& P = pa.view();
        LaneConst C;
        C.q0 = P.q0[ j ]; C.mass0 = P.mass0[ j ]; C.beta0 = P.beta0[ j ];
        C.p0c = P.p0c[ j ]; C.chi = P.chi[ j ]; C.charge_ratio = P.charge_ratio[ j ];
        L.delta = P.delta[ j ];
        cavity( L, C, e );
        if( active ) { P.delta[ j ] = L.delta; P.psigma[ j ] = L.psigma; }
        return true;
    }
    case ST_OP_XYSHIFT: xyshift( L, e ); return true;
    case ST_OP_SROTATION: srotation( L, e ); return true;
    case ST_OP_LIMIT_RECT: return limit_rect_ok( L, e );
    case ST_OP_LIMIT_ELLIPSE: return limit_ellipse_ok( L, e );
    case ST_OP_LIMIT_RECT_ELLIPSE: return limit_rect_ellipse_ok( L, e );
    case ST_OP_DIPEDGE: dipedge( L, e ); return true;
    default: break;
    }
    fail = true;
    return false;
}

// The same for the pair: one dispatch, each map applied to both particles
// (ok0 / ok1 their survival predicates).
template< class PA >
ST_DEV void lean_other_hot2( unsigned op, cptr_t e, Lane& L0, Lane& L1, const PA& pa, int64_t i0,
                             int64_t i1, bool active0, bool active1, bool& fail, bool& ok0, bool& ok1 )
{
    ok0 = ok1 = true;
    switch( op )
    {
    case ST_OP_XYSHIFT: xyshift( L0, e ); xyshift( L1, e ); return;
    case ST_OP_SROTATION: srotation( L0, e ); srotation( L1, e ); return;
    case ST_OP_LIMIT_RECT: ok0 = limit_rect_ok( L0, e ); ok1 = limit_rect_ok( L1, e ); return;
    case ST_OP_LIMIT_ELLIPSE: ok0 = limit_ellipse_ok( L0, e ); ok1 = limit_ellipse_ok( L1, e ); return;
    case ST_OP_LIMIT_RECT_ELLIPSE:
        ok0 = limit_rect_ellipse_ok( L0, e );
        ok1 = limit_rect_ellipse_ok( L1, e );
        return;
    case ST_OP_DIPEDGE: dipedge( L0, e ); dipedge( L1, e ); return;
    case ST_OP_DRIFT_EXACT:
    case ST_OP_CAVITY:
        ok0 = lean_other_hot( op, e, L0, pa, i0, active0, fail );
        ok1 = lean_other_hot( op, e, L1, pa, i1, active1, fail );
        return;
    default: break;
    }
    fail = true;
    ok0 = ok1 = false;
}

// Horner recurrence of two particles through one record: the coefficient
// reads and the order dispatch are shared, the two chains interleave.
template< int N >
ST_DEV void horner_fixed2( cptr_t e, double const x0, double const y0, double const x1,
                           double const y1, double& dpx0, double& dpy0, double& dpx1, double& dpy1 )
{
    dpx0 = dpx1 = rd( e, 2 * N );
    dpy0 = dpy1 = rd( e, 2 * N + 1 );
#pragma unroll
    for( int k = N; k > 0; --k )
    {
        double const bre = rd( e, 2 * k - 2 ), bim = rd( e, 2 * k - 1 );
        horner_step( dpx0, dpy0, x0, y0, bre, bim );
        horner_step( dpx1, dpy1, x1, y1, bre, bim );
    }
}

// be_multipole/track.h:38-112 for two particles without curvature terms
// (the caller runs multipole() for records with hxl or hyl != 0 and for
// orders above 10).
ST_DEV void multipole2_flat( Lane& p0, Lane& p1, double const chi0, double const chi1, cptr_t e,
                             int const order )
{
    double dpx0, dpy0, dpx1, dpy1;
    double const x0 = p0.x, y0 = p0.y, x1 = p1.x, y1 = p1.y;
    switch( order )
    {
    case 0: horner_fixed2< 0 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 1: horner_fixed2< 1 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 2: horner_fixed2< 2 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 3: horner_fixed2< 3 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 4: horner_fixed2< 4 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 5: horner_fixed2< 5 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 6: horner_fixed2< 6 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 7: horner_fixed2< 7 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 8: horner_fixed2< 8 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    case 9: horner_fixed2< 9 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    default: horner_fixed2< 10 >( e, x0, y0, x1, y1, dpx0, dpy0, dpx1, dpy1 ); break;
    }
#if defined( ST_FMA_BUILD )
    p0.px = fma( -chi0, dpx0, p0.px );
    p0.py = fma( chi0, dpy0, p0.py );
    p1.px = fma( -chi1, dpx1, p1.px );
    p1.py = fma( chi1, dpy1, p1.py );
#else
    dpx0 = -chi0 * dpx0;
    dpy0 = chi0 * dpy0;
    dpx1 = -chi1 * dpx1;
    dpy1 = chi1 * dpy1;
    p0.px += dpx0;
    p0.py += dpy0;
    p1.px += dpx1;
    p1.py += dpy1;
#endif
}

// Two particles per lane: lane l of a wave tracks the particles of slots
// base + l and base + 64 + l (list entries, or indices without a list)
// through the same records.  Each particle's maps, loss handling and stored
// values are those of track_lean_lane, bit for bit; what the pair shares is
// the record walk (scalar loads, header tests, branches, a multipole's order
// dispatch and coefficient reads), and the two Horner chains of a multipole
// interleave.  Only the state the Drift and Multipole records use stays in
// registers (x, px, y, py, zeta, rpp, rvv and chi: 16 VGPRs a particle), so
// that the pair fits the 64 VGPRs of eight waves per SIMD:
//   * delta and psigma stay in memory: Cavity and DriftExact read delta
//     there, Cavity writes both back, and a store leaves them as they are;
//   * every turn starts at s = +0 and s after a drift is the record's, as
//     in lean_turn's turns without a per-lane s;
//   * at_element and at_turn are read back at a store (memory holds their
//     values at the start of the launch until then); the turns a lane has
//     left are a 32-bit count against the wave's turn counter.
// A tile with a particle that starts its turn at s != 0 runs through
// track_lean_lane, one half after the other.
template< class PA >
ST_DEV void track_lean_lane2( const PA& pa, const int32_t* list, int64_t n_lanes,
                               int64_t base, cptr_t lat, int64_t n_elem, int64_t until_turn,
                               int64_t max_turns, int32_t* status_out, const WavePrio& wp = WavePrio() )
{
    constexpr int K = 2;
    int32_t idx[ K ], left[ K ];               // particle index, turns this launch
    double chi[ K ];
    Lane L[ K ];
    uint64_t act[ K ];
    bool mid = false;                          // a particle starts at s != 0
#pragma unroll
    for( int k = 0; k < K; ++k )
    {
        auto const& P = pa.view();
        int64_t const tid = base + 64 * k + __lane_id();
        bool const in_range = tid < n_lanes;
        int64_t const i = ( list != nullptr && in_range ) ? ( int64_t )list[ tid ] : tid;
        int64_t const li = in_range ? i : 0;
        idx[ k ] = ( int32_t )li;
        L[ k ].x = P.x[ li ]; L[ k ].px = P.px[ li ]; L[ k ].y = P.y[ li ];
        L[ k ].py = P.py[ li ]; L[ k ].zeta = P.zeta[ li ];
        L[ k ].rpp = P.rpp[ li ]; L[ k ].rvv = P.rvv[ li ];
        chi[ k ] = P.chi[ li ];
        int64_t const at_turn = P.at_turn[ li ];
        bool const active0 = in_range && n_elem > 0 && until_turn > at_turn && P.state[ li ] == 1;
        int64_t const stop = ( max_turns > 0 && until_turn - at_turn > max_turns ) ? at_turn + max_turns
                                                                                     : until_turn;
        left[ k ] = active0 ? ( int32_t )( stop - at_turn < INT32_MAX ? stop - at_turn : INT32_MAX ) : 0;
        mid = mid || ( active0 && __double_as_longlong( P.s[ li ] ) != 0 );
        act[ k ] = ballot( active0 );
    }
    if( ballot( mid ) != 0u || n_elem <= 0 )
    {
        track_lean_lane< false >( pa, list, n_lanes, base + __lane_id(), lat, n_elem,
                                  until_turn, max_turns, status_out, wp );
        track_lean_lane< false >( pa, list, n_lanes, base + 64 + __lane_id(), lat, n_elem,
                                  until_turn, max_turns, status_out, WavePrio() );
        return;
    }
    bool failed = false;                       // a lane met a disabled / unknown map
    int64_t turns_done = 0;                    // by this wave, in this call
    wave_priority( wp, 0 );
    while( ( act[ 0 ] | act[ 1 ] ) != 0u )
    {
        uint64_t s_rec = 0;                    // s after the last drift (bits)
        uint32_t off = 0;
        uint64_t h = rec_at( lat, 0 )[ 0 ];
        for( ;; )
        {
            while( ( ( uint32_t )h & ST_OP_RARE_BIT ) == 0u )   // Drift / Multipole run
            {
                cptr_t const r = rec_at( lat, off );
                uint32_t const noff = off + ST_REC_BYTES( h );
                uint64_t const hn = rec_at( lat, noff )[ 0 ];
                if( ( uint32_t )h & ST_OP_MP_BIT )
                {
                    unsigned const aux = ST_REC_AUX( h );
                    if( ( aux & ST_MP_CURV_FLAG ) == 0u && ( aux & 0x7fffu ) <= 10u )
                    {                          // the common record: one dispatch for both
                        multipole2_flat( L[ 0 ], L[ 1 ], chi[ 0 ], chi[ 1 ], r + ST_REC_PAYLOAD,
                                         ( int )( aux & 0x7fffu ) );
                    }
                    else if( aux & ST_MP_CURV_FLAG )
                    {
#pragma unroll
                        for( int k = 0; k < K; ++k )
                        {
                            int64_t j = idx[ k ];
                            asm volatile( "" : "+v"( j ) );
                            multipole( L[ k ], chi[ k ], r + ST_REC_PAYLOAD, aux, pa.view().delta[ j ] );
                        }
                    }
                    else
                    {
#pragma unroll
                        for( int k = 0; k < K; ++k )
                            multipole( L[ k ], chi[ k ], r + ST_REC_PAYLOAD, aux, 0.0 );
                    }
                }
                else
                {
                    double const length = rd( r, 2 );
                    s_rec = r[ 3 ];
#pragma unroll
                    for( int k = 0; k < K; ++k ) drift_xz( L[ k ], length );
                    uint64_t const lost0 = limit_global_lost_mask( L[ 0 ] ) & act[ 0 ];
                    uint64_t const lost1 = limit_global_lost_mask( L[ 1 ] ) & act[ 1 ];
                    if( ( lost0 | lost1 ) != 0u )
#pragma unroll
                    for( int k = 0; k < K; ++k )
                    {
                        uint64_t const lost = k == 0 ? lost0 : lost1;
                        if( lost != 0u )
                        {
                            if( lane_in( lost ) )
                            {
                                int64_t const j = idx[ k ];
                                auto const& P = pa.view();
                                store_hot_state( P, j, L[ k ], __longlong_as_double( ( long long )s_rec ), 0,
                                                 P.at_element[ j ] + ri( r, 1 ), P.at_turn[ j ] + turns_done );
                            }
                            act[ k ] &= ~lost;
                        }
                    }
                }
                off = noff;
                h = hn;
            }
            unsigned const op = ST_REC_OP( h );
            if( op == ST_OP_END ) break;
            cptr_t const r = rec_at( lat, off );
            uint32_t const noff = off + ST_REC_BYTES( h );
            uint64_t const hn = rec_at( lat, noff )[ 0 ];
            if( op == ST_OP_DRIFT_EXACT ) s_rec = r[ 3 ];
            bool fail = false, ok[ K ];
            lean_other_hot2( op, r + ST_REC_PAYLOAD, L[ 0 ], L[ 1 ], pa, idx[ 0 ], idx[ 1 ],
                             lane_in( act[ 0 ] ), lane_in( act[ 1 ] ), fail, ok[ 0 ], ok[ 1 ] );
#pragma unroll
            for( int k = 0; k < K; ++k )
            {
                uint64_t const lost = ballot( !ok[ k ] ) & act[ k ];
                if( lost != 0u )
                {
                    if( lane_in( lost ) )
                    {
                        int64_t const j = idx[ k ];
                        auto const& P = pa.view();
                        store_hot_state( P, j, L[ k ], __longlong_as_double( ( long long )s_rec ), 0,
                                         P.at_element[ j ] + ri( r, 1 ), P.at_turn[ j ] + turns_done );
                    }
                    // fail depends on the record alone: the same in every lane
                    failed = failed || fail;
                    act[ k ] &= ~lost;
                }
            }
            off = noff;
            h = hn;
        }
        ++turns_done;
#pragma unroll
        for( int k = 0; k < K; ++k )
        {
            uint64_t const fin = ballot( turns_done >= left[ k ] ) & act[ k ];
            if( fin != 0u )
            {
                if( lane_in( fin ) )
                {
                    int64_t const j = idx[ k ];
                    auto const& P = pa.view();
                    store_hot_state( P, j, L[ k ], 0.0, 1, P.at_element[ j ], P.at_turn[ j ] + turns_done );
                }
                act[ k ] &= ~fin;
            }
        }
        if( ( act[ 0 ] | act[ 1 ] ) != 0u ) wave_priority( wp, turns_done );
    }
    wave_priority_leave( wp, turns_done );
    if( failed && __lane_id() == 0 ) atomicOr( status_out, -1 );
}

// Indices of the particles a track_until call has work for (state 1 and
// at_turn < until_turn) appended, wave by wave, to ``list``: one atomic add
// per wave reserves the wave's slots, so the list is in no particular
// order (every particle is tracked independently; results do not depend
// on it).  counts[1] receives the smallest at_turn among them (clamped to
// int32).  Called by every lane of full waves.
ST_DEV void select_active_lane( const st_dev_particles_t& P, int64_t i, int64_t until_turn,
                                int32_t* list, int32_t* counts )
{
    int64_t const t = ( i < P.n ) ? P.at_turn[ i ] : 0;
    bool const want = i < P.n && P.state[ i ] == 1 && t < until_turn;
    uint64_t const m = ballot( want );
    if( m == 0u ) return;
    int32_t const t32 = ( int32_t )( want ? ( t < 0 ? 0 : ( t > INT32_MAX ? INT32_MAX : t ) )
                                          : INT32_MAX );
    int32_t tmin = t32;
    for( int off = 32; off > 0; off >>= 1 ) tmin = min( tmin, __shfl_xor( tmin, off ) );
    int base = 0;
    if( __lane_id() == 0 )
    {
        base = atomicAdd( &counts[ 0 ], __popcll( m ) );
        atomicMin( &counts[ 1 ], tmin );
    }
    base = __shfl( base, 0 );
    uint64_t const below = ( __lane_id() == 0 ) ? 0u : ( m << ( 64 - __lane_id() ) );
    if( want ) list[ base + __popcll( below ) ] = ( int32_t )i;
}

}  // namespace st

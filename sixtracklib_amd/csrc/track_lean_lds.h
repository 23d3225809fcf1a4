// Lean until-turn tracking with the lattice staged through LDS (device
// code), the alternative to the scalar-load record walk of track_lean.h:
// the waves of a workgroup track one tile of blockDim.x particles of one
// particle set, walk the lattice together, and copy it piece by piece
// (ST_LAT_STAGE_WORDS words) into LDS with coalesced vector loads, one copy
// per workgroup instead of one scalar-cache fill per wave.  Drift and
// Multipole records are read from LDS (every lane reads the same word, a
// broadcast); the other records of a lean lattice, a few per turn, are read
// from the stream as in track_lean.h.
//
// Results are those of track_lean_lane bit for bit: the same maps in the
// same order with the same roundings (the FMA build's Horner step with the
// coefficient as an ordinary operand instead of an SGPR addend is the same
// fused operation).  Lean lattices only; one particle per lane.
#pragma once

#include "track_lean.h"

namespace st
{
typedef __attribute__( ( address_space( 3 ) ) ) const uint64_t lword_t;
typedef lword_t* lptr_t;

ST_DEV double rdl( lptr_t p, int k ) { return __longlong_as_double( ( long long )p[ k ] ); }

ST_DEV void horner_step_v( double& dpx, double& dpy, double const x, double const y,
                           double const bre, double const bim )
{
#if defined( ST_FMA_BUILD )
    double const ndpx = fma( dpx, x, fma( -dpy, y, bre ) );
    dpy = fma( dpx, y, fma( dpy, x, bim ) );
    dpx = ndpx;
#else
    double const zre = dpx * x - dpy * y;
    double const zim = dpx * y + dpy * x;
    dpx = bre + zre;
    dpy = bim + zim;
#endif
}

template< int N >
ST_DEV void horner_fixed_l( lptr_t e, double const x, double const y, double& dpx, double& dpy )
{
    dpx = rdl( e, 2 * N );
    dpy = rdl( e, 2 * N + 1 );
#pragma unroll
    for( int k = N; k > 0; --k )
        horner_step_v( dpx, dpy, x, y, rdl( e, 2 * k - 2 ), rdl( e, 2 * k - 1 ) );
}

ST_DEV void horner_l( lptr_t e, int const order, double const x, double const y, double& dpx,
                      double& dpy )
{
    switch( order )
    {
    case 0: horner_fixed_l< 0 >( e, x, y, dpx, dpy ); return;
    case 1: horner_fixed_l< 1 >( e, x, y, dpx, dpy ); return;
    case 2: horner_fixed_l< 2 >( e, x, y, dpx, dpy ); return;
    case 3: horner_fixed_l< 3 >( e, x, y, dpx, dpy ); return;
    case 4: horner_fixed_l< 4 >( e, x, y, dpx, dpy ); return;
    case 5: horner_fixed_l< 5 >( e, x, y, dpx, dpy ); return;
    case 6: horner_fixed_l< 6 >( e, x, y, dpx, dpy ); return;
    case 7: horner_fixed_l< 7 >( e, x, y, dpx, dpy ); return;
    case 8: horner_fixed_l< 8 >( e, x, y, dpx, dpy ); return;
    case 9: horner_fixed_l< 9 >( e, x, y, dpx, dpy ); return;
    case 10: horner_fixed_l< 10 >( e, x, y, dpx, dpy ); return;
    default: break;
    }
    int ix = 2 * order;
    dpx = rdl( e, ix );
    dpy = rdl( e, ix + 1 );
    for( ; ix > 0; ix -= 2 ) horner_step_v( dpx, dpy, x, y, rdl( e, ix - 2 ), rdl( e, ix - 1 ) );
}

// multipole() of track_device.h (be_multipole/track.h:38-112) on a record in LDS
ST_DEV void multipole_l( Lane& p, double const chi, lptr_t e, unsigned aux, double const delta )
{
    int const order = ( int )( aux & 0x7fffu );
    double dpx, dpy;
    double const x = p.x, y = p.y;
    horner_l( e, order, x, y, dpx, dpy );
#if defined( ST_FMA_BUILD )
    if( !( aux & ST_MP_CURV_FLAG ) )
    {
        p.px = fma( -chi, dpx, p.px );
        p.py = fma( chi, dpy, p.py );
        return;
    }
#endif
    dpx = -chi * dpx;
    dpy = chi * dpy;
    if( aux & ST_MP_CURV_FLAG )
    {
        int const nb = 2 * order + 2;
        double const hxl = rdl( e, nb ), hyl = rdl( e, nb + 1 ), length = rdl( e, nb + 2 );
        double const hxlx = x * hxl;
        double const hyly = y * hyl;
        p.zeta += chi * ( hyly - hxlx );
        dpx += hxl + hxl * delta;
        dpy -= hyl + hyl * delta;
        if( length > 0.0 )
        {
            double const b1l = chi * rdl( e, 0 );
            double const a1l = chi * rdl( e, 1 );
            dpx -= b1l * hxlx / length;
            dpy += a1l * hyly / length;
        }
    }
    p.px += dpx;
    p.py += dpy;
}

// A wave-uniform 32-bit value read from LDS as a scalar.
ST_DEV uint32_t uniform_u32( uint32_t v ) { return ( uint32_t )__builtin_amdgcn_readfirstlane( ( int )v ); }

// One workgroup of a lean until-turn launch with LDS staging: particle
// tile_base + threadIdx.x of the set (list entries, or indices without a
// list; n_lanes valid), the whole workgroup on one lattice.  ``stage`` is
// the workgroup's LDS buffer of ST_LAT_STAGE_WORDS words.  The OR of the
// per-particle statuses goes to *status_out.
ST_DEV void track_lean_block( const st_dev_particles_t& P, const int32_t* list, int64_t n_lanes,
                              int64_t tile_base, cptr_t lat, int64_t n_elem, int64_t until_turn,
                              int64_t max_turns, int32_t* status_out,
                              __attribute__( ( address_space( 3 ) ) ) uint64_t* stage )
{
    int64_t const tid = tile_base + threadIdx.x;
    bool const in_range = tid < n_lanes;
    int64_t const i = ( list != nullptr && in_range ) ? ( int64_t )list[ tid ] : tid;
    int64_t const li = in_range ? i : 0;
    PDirect const pa{ P };
    Lane L;
    L.x = P.x[ li ]; L.px = P.px[ li ]; L.y = P.y[ li ]; L.py = P.py[ li ];
    L.zeta = P.zeta[ li ]; L.delta = P.delta[ li ]; L.rpp = P.rpp[ li ];
    L.rvv = P.rvv[ li ]; L.psigma = P.psigma[ li ]; L.s = P.s[ li ];
    double const chi = P.chi[ li ];
    int64_t const start = P.at_element[ li ];
    int64_t at_turn = P.at_turn[ li ];
    bool const active0 = in_range && n_elem > 0 && until_turn > at_turn && P.state[ li ] == 1;
    int32_t status = ( in_range && n_elem <= 0 ) ? -1 : 0;
    int64_t const turn_stop = ( max_turns > 0 && until_turn - at_turn > max_turns )
                                  ? at_turn + max_turns : until_turn;
    uint64_t act = ballot( active0 );
    lptr_t const st = ( lptr_t )stage;
    // the workgroup tracks turns while any of its waves has active lanes
    while( __syncthreads_or( act != 0u ) )
    {
        // every lane of the workgroup walks the records; a wave without
        // active lanes skips the maps (their results are never stored)
        bool const track = act != 0u;
        uint32_t base = 0;                    // stream word of stage[ 0 ]
        bool done = false;
        while( !done )
        {
            __syncthreads();                  // the previous piece is no longer read
            {
                const uint2* src = ( const uint2* )( ( const uint64_t* )lat + base );
                for( int k = threadIdx.x; k < ST_LAT_STAGE_WORDS; k += blockDim.x )
                {
                    uint2 const v = src[ k ];
                    stage[ k ] = ( uint64_t )v.x | ( ( uint64_t )v.y << 32 );
                }
            }
            __syncthreads();
            // records that lie in this piece with the next record's header:
            // from LDS; a record beyond it starts the next piece
            uint32_t lo = 0;
            for( ;; )
            {
                uint64_t const h = st[ lo ];
                uint32_t const hlo = uniform_u32( ( uint32_t )h );
                uint32_t const bytes = uniform_u32( ( uint32_t )( h >> 32 ) );
                uint32_t const op = hlo & 0xffu;
                if( op == ST_OP_END ) { done = true; break; }
                uint32_t const nlo = lo + ( bytes >> 3 );
                if( nlo + 2u > ( uint32_t )ST_LAT_STAGE_WORDS )
                {
                    if( lo == 0u )
                    {
                        // a record longer than the stage (the host does not
                        // stage such lattices): every lane stops, status -1
                        if( in_range ) status = -1;
                        act = 0u;
                        done = true;
                        break;
                    }
                    base += lo;
                    break;
                }
                if( ( hlo & ST_OP_RARE_BIT ) == 0u )
                {
                    if( track )
                    {
                        lptr_t const r = st + lo;
                        if( hlo & ST_OP_MP_BIT )
                            multipole_l( L, chi, r + ST_REC_PAYLOAD, ( hlo >> 8 ) & 0xffffu, L.delta );
                        else
                        {
                            drift( L, rdl( r, ST_REC_PAYLOAD ) );
                            uint64_t const lost = limit_global_lost_mask( L ) & act;
                            if( lost != 0u )
                            {
                                if( lane_in( lost ) )
                                    store_lane_state( P, i, L, 0, start + ( int64_t )r[ 1 ], at_turn );
                                act &= ~lost;
                            }
                        }
                    }
                }
                else if( track )
                {
                    // other lean maps: the record from the stream itself
                    cptr_t const r = rec_at( lat, ( base + lo ) * 8u );
                    bool fail = false;
                    bool const ok = lean_other< false >( op, ( hlo >> 8 ) & 0xffffu, r + ST_REC_PAYLOAD,
                                                         L, pa, li, lane_in( act ), start + ri( r, 1 ),
                                                         at_turn, fail );
                    uint64_t const lost = ballot( !ok ) & act;
                    if( lost != 0u )
                    {
                        if( lane_in( lost ) )
                        {
                            store_lane_state( P, i, L, 0, start + ri( r, 1 ), at_turn );
                            if( fail ) status = -1;
                        }
                        act &= ~lost;
                    }
                }
                lo = nlo;
            }
        }
        if( track )
        {
            if( lane_in( act ) ) { ++at_turn; L.s = 0.0; }
            uint64_t const fin = ballot( at_turn >= turn_stop ) & act;
            if( fin != 0u )
            {
                if( lane_in( fin ) ) store_lane_state( P, i, L, 1, start, at_turn );
                act &= ~fin;
            }
        }
    }
    if( in_range && status != 0 ) atomicOr( status_out, status );
}

}  // namespace st

// Records of whole particles (element-by-element output, BeamMonitor
// stores) and the heavy maps as seen from the lean until-turn loop
// (device code).
#pragma once

#include "track_device.h"

namespace st
{

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

// device copy of an output particle set header, addressed by a monitor
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

// be_monitor/track.h:47-170.  Returns true when the copy fails (index past
// the output set): the reference then marks the particle lost, status -1.
ST_DEV_NOINLINE bool monitor( cptr_t e, const st_dev_particles_t& P, int64_t i,
    const Lane& L, const LaneConst& C, int64_t pid, int64_t at_element,
    int64_t at_turn, bool active )
{
    uint64_t const out = e[ 3 ];
    if( out == 0u ) return false;
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
    if( out_idx < 0 ) return false;
    st_dev_particles_t const O = load_set( out );
    if( out_idx >= O.n ) return true;
    if( active ) store_record( O, out_idx, P, i, L, C, pid, at_element, at_turn );
    return false;
}

// The maps outside the lean set, applied to one lane from the lean
// until-turn loop (track/track.h:209-585 for these types).  Out of line so
// that their register needs do not set the lean loop's budget; the lane
// state travels through ``io`` and the per-particle constants are read
// here.  ``active``: the lane is still tracked (only then does a monitor
// store).  Returns the survival predicate; ``fail`` is set when the map
// reports failure (a monitor whose index lies past its output set, a
// disabled map, an unknown type).
ST_DEV_NOINLINE bool heavy_step( unsigned op, unsigned aux, cptr_t e, Lane* io,
    cptr_t set_desc, int64_t li, bool active, int64_t at_element,
    int64_t at_turn, bool* fail )
{
    st_dev_particles_t const PV = load_set( ( uint64_t )set_desc );
    st_dev_particles_t const* const P = &PV;
    int64_t j = li;
    asm volatile( "" : "+v"( j ) );
    LaneConst C;
    C.q0 = P->q0[ j ]; C.mass0 = P->mass0[ j ]; C.beta0 = P->beta0[ j ];
    C.p0c = P->p0c[ j ]; C.chi = P->chi[ j ]; C.charge_ratio = P->charge_ratio[ j ];
    Lane& L = *io;
    switch( op )
    {
    case ST_OP_MONITOR:
    {
        L.state = 1;                              // a tracked lane is alive
        int64_t const pid = P->particle_id[ j ];
        if( monitor( e, *P, j, L, C, pid, at_element, at_turn, active ) )
        { *fail = true; return false; }
        return true;
    }
    case ST_OP_BB4D: bb4d( L, C, e ); return true;
    case ST_OP_BB6D: bb6d( L, C, e ); return true;
    case ST_OP_SC_COASTING: sc_coasting( L, C, e ); return true;
    case ST_OP_SC_QGAUSS: sc_qgauss( L, C, e ); return true;
    case ST_OP_SC_INTERPOL: sc_interpol( L, C, e ); return true;
    case ST_OP_RF_MULTIPOLE: rf_multipole( L, C, e, aux ); return true;
    default: break;
    }
    *fail = true;                                 // disabled map or unknown type
    return false;
}

}  // namespace st

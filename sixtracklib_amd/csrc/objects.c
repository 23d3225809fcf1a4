/* Particles and beam-element builders on CObjects buffers (host, C99). */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "sixtracklib_amd/particles.h"
#include "sixtracklib_amd/beam_elements.h"

/* ---------------------------------------------------------------- Particles */

NS(Particles)* NS(Particles_new)( NS(Buffer)* b, NS(particle_num_elements_t) n )
{
    NS(Particles) tmpl;
    NS(buffer_size_t) offsets[ 21 ], sizes[ 21 ], counts[ 21 ];
    NS(Object)* obj;
    int k;
    if( b == NULL || n < 0 ) return NULL;
    memset( &tmpl, 0, sizeof( tmpl ) );
    tmpl.num_particles = n;
    for( k = 0 ; k < 21 ; ++k )
    {
        offsets[ k ] = 8u + 8u * ( NS(buffer_size_t) )k;
        sizes[ k ] = 8u;
        counts[ k ] = ( NS(buffer_size_t) )n;
    }
    obj = NS(Buffer_add_object)( b, &tmpl, sizeof( tmpl ),
        ST_OBJECT_TYPE_PARTICLE, 21u, offsets, sizes, counts );
    return ( obj != NULL ) ? ( NS(Particles)* )( uintptr_t )obj->begin_addr : NULL;
}

NS(Particles)* NS(Particles_add_copy)( NS(Buffer)* b, const NS(Particles)* orig )
{
    NS(Particles)* p;
    NS(Particles) snapshot;
    if( orig == NULL ) return NULL;
    snapshot = *orig;          /* orig may live in b and move on growth */
    p = NS(Particles_new)( b, snapshot.num_particles );
    if( p != NULL ) NS(Particles_copy)( p, &snapshot );
    return p;
}

NS(Particles)* NS(Particles_buffer_get_particles)( NS(Buffer)* b,
                                                   NS(buffer_size_t) index )
{
    NS(Object)* obj = NS(Buffer_get_object)( b, index );
    if( obj == NULL || obj->type_id != ST_OBJECT_TYPE_PARTICLE ) return NULL;
    return ( NS(Particles)* )( uintptr_t )obj->begin_addr;
}

NS(Particles) const* NS(Particles_buffer_get_const_particles)(
    const NS(Buffer)* b, NS(buffer_size_t) index )
{
    return NS(Particles_buffer_get_particles)( ( NS(Buffer)* )b, index );
}

NS(particle_num_elements_t) NS(Particles_get_num_of_particles)( const NS(Particles)* p )
{
    return ( p != NULL ) ? p->num_particles : 0;
}

NS(buffer_size_t) NS(Particles_buffer_get_num_of_particle_blocks)( const NS(Buffer)* b )
{
    NS(buffer_size_t) i, n = NS(Buffer_get_num_of_objects)( b ), cnt = 0u;
    for( i = 0 ; i < n ; ++i )
        if( NS(Buffer_get_const_object)( b, i )->type_id == ST_OBJECT_TYPE_PARTICLE ) ++cnt;
    return cnt;
}

NS(particle_num_elements_t) NS(Particles_buffer_get_total_num_of_particles)(
    const NS(Buffer)* b )
{
    NS(buffer_size_t) i, n = NS(Buffer_get_num_of_objects)( b );
    NS(particle_num_elements_t) total = 0;
    for( i = 0 ; i < n ; ++i )
    {
        NS(Particles) const* p = NS(Particles_buffer_get_const_particles)( b, i );
        if( p != NULL ) total += p->num_particles;
    }
    return total;
}

int NS(Particles_copy)( NS(Particles)* d, const NS(Particles)* s )
{
    int k;
    size_t bytes;
    if( d == NULL || s == NULL || d->num_particles != s->num_particles ) return -1;
    bytes = ( size_t )s->num_particles * 8u;
    for( k = 0 ; k < 21 ; ++k )
    {
        void* dst = ( ( void** )&d->q0 )[ k ];
        const void* src = ( ( void* const* )&s->q0 )[ k ];
        if( bytes ) memcpy( dst, src, bytes );
    }
    return 0;
}

int NS(Particles_find_min_max_attributes)( const NS(Particles)* p,
    NS(particle_index_t)* min_pid, NS(particle_index_t)* max_pid,
    NS(particle_index_t)* min_elem, NS(particle_index_t)* max_elem,
    NS(particle_index_t)* min_turn, NS(particle_index_t)* max_turn )
{
    NS(particle_num_elements_t) i;
    if( p == NULL ) return -1;
    for( i = 0 ; i < p->num_particles ; ++i )
    {
        int64_t const pid = p->particle_id[ i ];
        if( min_pid && pid < *min_pid ) *min_pid = pid;
        if( max_pid && pid > *max_pid ) *max_pid = pid;
        if( min_elem && p->at_element_id[ i ] < *min_elem ) *min_elem = p->at_element_id[ i ];
        if( max_elem && p->at_element_id[ i ] > *max_elem ) *max_elem = p->at_element_id[ i ];
        if( min_turn && p->at_turn[ i ] < *min_turn ) *min_turn = p->at_turn[ i ];
        if( max_turn && p->at_turn[ i ] > *max_turn ) *max_turn = p->at_turn[ i ];
    }
    return 0;
}

/* ------------------------------------------------------------- elements */

static void* add_pod( NS(Buffer)* b, const void* pod, size_t size,
                      NS(object_type_id_t) type_id )
{
    NS(Object)* obj = NS(Buffer_add_object)( b, pod, size, type_id, 0u,
                                             NULL, NULL, NULL );
    return ( obj != NULL ) ? ( void* )( uintptr_t )obj->begin_addr : NULL;
}

NS(Drift)* NS(Drift_add)( NS(Buffer)* b, double length )
{
    NS(Drift) e; e.length = length;
    return ( NS(Drift)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_DRIFT );
}

NS(DriftExact)* NS(DriftExact_add)( NS(Buffer)* b, double length )
{
    NS(DriftExact) e; e.length = length;
    return ( NS(DriftExact)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_DRIFT_EXACT );
}

NS(Multipole)* NS(Multipole_add)( NS(Buffer)* b, int64_t order, double length,
    double hxl, double hyl, NS(buffer_addr_t) bal_addr )
{
    NS(Multipole) e;
    NS(buffer_size_t) off = 32u, sz = 8u, cnt;
    NS(Object)* obj;
    if( order < 0 ) return NULL;
    e.order = order; e.length = length; e.hxl = hxl; e.hyl = hyl;
    e.bal_addr = bal_addr;
    cnt = ( NS(buffer_size_t) )( 2 * order + 2 );
    obj = NS(Buffer_add_object)( b, &e, sizeof( e ), ST_OBJECT_TYPE_MULTIPOLE,
                                 1u, &off, &sz, &cnt );
    return ( obj != NULL ) ? ( NS(Multipole)* )( uintptr_t )obj->begin_addr : NULL;
}

NS(Multipole)* NS(Multipole_new)( NS(Buffer)* b, int64_t order )
{
    return NS(Multipole_add)( b, order, 0.0, 0.0, 0.0, 0u );
}

NS(RFMultipole)* NS(RFMultipole_add)( NS(Buffer)* b, int64_t order,
    double voltage, double frequency, double lag, NS(buffer_addr_t) bal_addr,
    NS(buffer_addr_t) phase_addr )
{
    NS(RFMultipole) e;
    NS(buffer_size_t) off[ 2 ] = { 32u, 40u }, sz[ 2 ] = { 8u, 8u }, cnt[ 2 ];
    NS(Object)* obj;
    if( order < 0 ) return NULL;
    e.order = order; e.voltage = voltage; e.frequency = frequency; e.lag = lag;
    e.bal_addr = bal_addr; e.phase_addr = phase_addr;
    cnt[ 0 ] = cnt[ 1 ] = ( NS(buffer_size_t) )( 2 * order + 2 );
    obj = NS(Buffer_add_object)( b, &e, sizeof( e ), ST_OBJECT_TYPE_RF_MULTIPOLE,
                                 2u, off, sz, cnt );
    return ( obj != NULL ) ? ( NS(RFMultipole)* )( uintptr_t )obj->begin_addr : NULL;
}

NS(Cavity)* NS(Cavity_add)( NS(Buffer)* b, double v, double f, double lag )
{
    NS(Cavity) e; e.voltage = v; e.frequency = f; e.lag = lag;
    return ( NS(Cavity)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_CAVITY );
}

NS(XYShift)* NS(XYShift_add)( NS(Buffer)* b, double dx, double dy )
{
    NS(XYShift) e; e.dx = dx; e.dy = dy;
    return ( NS(XYShift)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_XYSHIFT );
}

NS(SRotation)* NS(SRotation_add)( NS(Buffer)* b, double c, double s )
{
    NS(SRotation) e; e.cos_z = c; e.sin_z = s;
    return ( NS(SRotation)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_SROTATION );
}

NS(LimitRect)* NS(LimitRect_add)( NS(Buffer)* b, double min_x, double max_x,
                                  double min_y, double max_y )
{
    NS(LimitRect) e; e.min_x = min_x; e.max_x = max_x; e.min_y = min_y; e.max_y = max_y;
    return ( NS(LimitRect)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_LIMIT_RECT );
}

NS(LimitEllipse)* NS(LimitEllipse_add)( NS(Buffer)* b, double a, double bb )
{
    NS(LimitEllipse) e;
    e.a_squ = a * a; e.b_squ = bb * bb; e.a_b_squ = e.a_squ * e.b_squ;
    return ( NS(LimitEllipse)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_LIMIT_ELLIPSE );
}

NS(LimitRectEllipse)* NS(LimitRectEllipse_add)( NS(Buffer)* b, double max_x,
    double max_y, double a_squ, double b_squ )
{
    NS(LimitRectEllipse) e;
    e.max_x = max_x; e.max_y = max_y; e.a_squ = a_squ; e.b_squ = b_squ;
    e.a_squ_b_squ = a_squ * b_squ;
    return ( NS(LimitRectEllipse)* )add_pod( b, &e, sizeof( e ),
                                             ST_OBJECT_TYPE_LIMIT_RECT_ELLIPSE );
}

NS(DipoleEdge)* NS(DipoleEdge_add)( NS(Buffer)* b, double r21, double r43 )
{
    NS(DipoleEdge) e; e.r21 = r21; e.r43 = r43;
    return ( NS(DipoleEdge)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_DIPEDGE );
}

NS(BeamMonitor)* NS(BeamMonitor_add)( NS(Buffer)* b, int64_t num_stores,
    int64_t start, int64_t skip, NS(buffer_addr_t) out_address,
    int64_t min_pid, int64_t max_pid, bool is_rolling, bool is_turn_ordered )
{
    NS(BeamMonitor) e;
    e.num_stores = num_stores; e.start = start; e.skip = skip;
    e.out_address = out_address; e.max_particle_id = max_pid;
    e.min_particle_id = min_pid; e.is_rolling = is_rolling ? 1 : 0;
    e.is_turn_ordered = is_turn_ordered ? 1 : 0;
    return ( NS(BeamMonitor)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_BEAM_MONITOR );
}

static void* add_bb( NS(Buffer)* b, const double* data, uint64_t n,
                     NS(object_type_id_t) tid )
{
    NS(BeamBeam4D) e;
    NS(buffer_size_t) off = 8u, sz = 8u, cnt = n;
    NS(Object)* obj;
    e.data_size = n;
    e.data_addr = ( NS(buffer_addr_t) )( uintptr_t )data;
    obj = NS(Buffer_add_object)( b, &e, sizeof( e ), tid, 1u, &off, &sz, &cnt );
    return ( obj != NULL ) ? ( void* )( uintptr_t )obj->begin_addr : NULL;
}

NS(BeamBeam4D)* NS(BeamBeam4D_add)( NS(Buffer)* b, const double* d, uint64_t n )
{
    return ( NS(BeamBeam4D)* )add_bb( b, d, n, ST_OBJECT_TYPE_BEAM_BEAM_4D );
}

NS(BeamBeam6D)* NS(BeamBeam6D_add)( NS(Buffer)* b, const double* d, uint64_t n )
{
    return ( NS(BeamBeam6D)* )add_bb( b, d, n, ST_OBJECT_TYPE_BEAM_BEAM_6D );
}

NS(SCCoasting)* NS(SCCoasting_add)( NS(Buffer)* b, double npart, double circ,
    double sx, double sy, double length, double x_co, double y_co,
    double min_sigma_diff, bool enabled )
{
    NS(SCCoasting) e;
    e.number_of_particles = npart; e.circumference = circ; e.sigma_x = sx;
    e.sigma_y = sy; e.length = length; e.x_co = x_co; e.y_co = y_co;
    e.min_sigma_diff = min_sigma_diff; e.enabled = enabled ? 1u : 0u;
    return ( NS(SCCoasting)* )add_pod( b, &e, sizeof( e ), ST_OBJECT_TYPE_SC_COASTING );
}

double NS(Math_q_gauss_cq)( double q )
{
    double const eps = 1e-6;
    double cq = ST_SQRT_PI;
    if( q >= 1.0 + eps )
    {
        double const qm1 = q - 1.0;
        cq *= tgamma( ( 3.0 - q ) / ( 2.0 * qm1 ) );
        cq /= sqrt( qm1 ) * tgamma( 1.0 / qm1 );
    }
    else if( q <= 1.0 - eps )
    {
        double const omq = 1.0 - q, tmq = 3.0 - q;
        cq *= 2.0 * tgamma( 1.0 / omq );
        cq /= tmq * sqrt( omq ) * tgamma( tmq / ( 2.0 * omq ) );
    }
    return cq;
}

NS(SCQGaussProfile)* NS(SCQGaussProfile_add)( NS(Buffer)* b, double npart,
    double bl_rms, double sx, double sy, double length, double x_co,
    double y_co, double min_sigma_diff, double q_param, bool enabled )
{
    NS(SCQGaussProfile) e;
    e.number_of_particles = npart; e.bunchlength_rms = bl_rms; e.sigma_x = sx;
    e.sigma_y = sy; e.length = length; e.x_co = x_co; e.y_co = y_co;
    e.min_sigma_diff = min_sigma_diff; e.q_param = q_param;
    e.cq = NS(Math_q_gauss_cq)( q_param ); e.enabled = enabled ? 1u : 0u;
    return ( NS(SCQGaussProfile)* )add_pod( b, &e, sizeof( e ),
                                            ST_OBJECT_TYPE_SC_QGAUSSIAN_PROF );
}

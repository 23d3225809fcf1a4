/* sixtracklib_amd/beam_elements.h -- beam-element PODs and builders.
 *
 * Struct layouts are identical to the reference element structs (cited on
 * each), so lattices are byte-compatible.  Builders append an element to
 * a CObjects buffer and return a pointer to it (valid until the buffer
 * grows), with the reference's argument order.
 */
#ifndef SIXTRACKLIB_AMD_BEAM_ELEMENTS_H__
#define SIXTRACKLIB_AMD_BEAM_ELEMENTS_H__

#include "sixtracklib_amd/definitions.h"
#include "sixtracklib_amd/buffer.h"

#if defined( __cplusplus )
extern "C" {
#endif

typedef struct NS(Drift) { double length; } NS(Drift);            /* be_drift/be_drift.h:28-31 */
typedef struct NS(DriftExact) { double length; } NS(DriftExact);  /* be_drift/be_drift.h:159-162 */
typedef struct NS(Multipole) { int64_t order; double length, hxl, hyl;
    NS(buffer_addr_t) bal_addr; } NS(Multipole);                /* be_multipole/be_multipole.h:27-34 */
typedef struct NS(RFMultipole) { int64_t order; double voltage, frequency,
    lag; NS(buffer_addr_t) bal_addr, phase_addr; } NS(RFMultipole); /* be_rfmultipole/be_rfmultipole.h */
typedef struct NS(Cavity) { double voltage, frequency, lag; } NS(Cavity); /* be_cavity/be_cavity.h:27-32 */
typedef struct NS(XYShift) { double dx, dy; } NS(XYShift);        /* be_xyshift/be_xyshift.h:26-30 */
typedef struct NS(SRotation) { double cos_z, sin_z; } NS(SRotation); /* be_srotation/be_srotation.h:26-30 */
typedef struct NS(LimitRect) { double min_x, max_x, min_y, max_y; }
    NS(LimitRect);                                              /* be_limit/be_limit_rect.h:26-33 */
typedef struct NS(LimitEllipse) { double a_squ, b_squ, a_b_squ; }
    NS(LimitEllipse);                                           /* be_limit/be_limit_ellipse.h:28-33 */
typedef struct NS(LimitRectEllipse) { double max_x, max_y, a_squ, b_squ,
    a_squ_b_squ; } NS(LimitRectEllipse);                        /* be_limit/be_limit_rect_ellipse.h:26-33 */
typedef struct NS(DipoleEdge) { double r21, r43; } NS(DipoleEdge); /* be_dipedge/be_dipedge.h:29-33 */
typedef struct NS(BeamMonitor) { int64_t num_stores, start, skip;
    NS(buffer_addr_t) out_address; int64_t max_particle_id, min_particle_id,
    is_rolling, is_turn_ordered; } NS(BeamMonitor);             /* be_monitor/be_monitor.h:35-45 */
typedef struct NS(BeamBeam4D) { uint64_t data_size; NS(buffer_addr_t) data_addr; }
    NS(BeamBeam4D);                                             /* be_beamfields/be_beamfields.h:36-41 */
typedef struct NS(BeamBeam6D) { uint64_t data_size; NS(buffer_addr_t) data_addr; }
    NS(BeamBeam6D);                                             /* be_beamfields/be_beamfields.h:194-199 */
typedef struct NS(SCCoasting) { double number_of_particles, circumference,
    sigma_x, sigma_y, length, x_co, y_co, min_sigma_diff; uint64_t enabled; }
    NS(SCCoasting);                                             /* be_beamfields/be_beamfields.h:418-429 */
typedef struct NS(SCQGaussProfile) { double number_of_particles,
    bunchlength_rms, sigma_x, sigma_y, length, x_co, y_co, min_sigma_diff,
    q_param, cq; uint64_t enabled; } NS(SCQGaussProfile);       /* be_beamfields/be_beamfields.h:638-651 */
typedef struct NS(SCInterpolatedProfile) { double number_of_particles,
    sigma_x, sigma_y, length, x_co, y_co; NS(buffer_addr_t) interpol_data_addr;
    double line_density_prof_fallback, min_sigma_diff; uint64_t enabled; }
    NS(SCInterpolatedProfile);                                  /* be_beamfields/be_beamfields.h:1281-1293 */
typedef struct NS(LineDensityProfileData) { int64_t method, capacity;
    NS(buffer_addr_t) values_addr, derivatives_addr; double z0, dz;
    int64_t num_values; } NS(LineDensityProfileData);           /* be_beamfields/be_beamfields.h:887-896 */

/* be_drift/be_drift.h:147 */
ST_API NS(Drift)* NS(Drift_add)( NS(Buffer)* b, double length );
/* be_drift/be_drift.h:280 */
ST_API NS(DriftExact)* NS(DriftExact_add)( NS(Buffer)* b, double length );
/* be_multipole/be_multipole.h:271 (bal_addr 0 -> zero coefficients) */
ST_API NS(Multipole)* NS(Multipole_add)( NS(Buffer)* b, int64_t order,
    double length, double hxl, double hyl, NS(buffer_addr_t) bal_addr );
/* be_multipole/be_multipole.h: NS(Multipole_new) */
ST_API NS(Multipole)* NS(Multipole_new)( NS(Buffer)* b, int64_t order );
/* be_rfmultipole/be_rfmultipole.h:357 */
ST_API NS(RFMultipole)* NS(RFMultipole_add)( NS(Buffer)* b, int64_t order,
    double voltage, double frequency, double lag,
    NS(buffer_addr_t) bal_addr, NS(buffer_addr_t) phase_addr );
/* be_cavity/be_cavity.h:155 */
ST_API NS(Cavity)* NS(Cavity_add)( NS(Buffer)* b, double voltage,
    double frequency, double lag );
/* be_xyshift/be_xyshift.h:154 */
ST_API NS(XYShift)* NS(XYShift_add)( NS(Buffer)* b, double dx, double dy );
/* be_srotation/be_srotation.h:170 */
ST_API NS(SRotation)* NS(SRotation_add)( NS(Buffer)* b, double cos_angle,
    double sin_angle );
/* be_limit/be_limit_rect.h:181 */
ST_API NS(LimitRect)* NS(LimitRect_add)( NS(Buffer)* b, double min_x,
    double max_x, double min_y, double max_y );
/* be_limit/be_limit_ellipse.h:174 */
ST_API NS(LimitEllipse)* NS(LimitEllipse_add)( NS(Buffer)* b,
    double x_semi_axis, double y_semi_axis );
/* be_limit/be_limit_rect_ellipse.h:199 */
ST_API NS(LimitRectEllipse)* NS(LimitRectEllipse_add)( NS(Buffer)* b,
    double max_x, double max_y, double a_squ, double b_squ );
/* be_dipedge/be_dipedge.h:164 */
ST_API NS(DipoleEdge)* NS(DipoleEdge_add)( NS(Buffer)* b, double r21, double r43 );
/* be_monitor/be_monitor.h:419 */
ST_API NS(BeamMonitor)* NS(BeamMonitor_add)( NS(Buffer)* b, int64_t num_stores,
    int64_t start, int64_t skip, NS(buffer_addr_t) out_address,
    int64_t min_particle_id, int64_t max_particle_id, bool is_rolling,
    bool is_turn_ordered );
/* be_beamfields/be_beamfields.h: NS(BeamBeam4D_add) / NS(BeamBeam6D_add)
 * (data copied from ``data``, ``data_size`` 8-byte words) */
ST_API NS(BeamBeam4D)* NS(BeamBeam4D_add)( NS(Buffer)* b, const double* data,
    uint64_t data_size );
ST_API NS(BeamBeam6D)* NS(BeamBeam6D_add)( NS(Buffer)* b, const double* data,
    uint64_t data_size );
/* be_beamfields/be_beamfields.h: NS(SCCoasting_add) */
ST_API NS(SCCoasting)* NS(SCCoasting_add)( NS(Buffer)* b,
    double number_of_particles, double circumference, double sigma_x,
    double sigma_y, double length, double x_co, double y_co,
    double min_sigma_diff, bool enabled );
/* be_beamfields/be_beamfields.h: NS(SCQGaussProfile_add) */
ST_API NS(SCQGaussProfile)* NS(SCQGaussProfile_add)( NS(Buffer)* b,
    double number_of_particles, double bunchlength_rms, double sigma_x,
    double sigma_y, double length, double x_co, double y_co,
    double min_sigma_diff, double q_param, bool enabled );

/* internal/math_qgauss.h:137-159 NS(Math_q_gauss_cq) (host) */
ST_API double NS(Math_q_gauss_cq)( double q );

#if defined( __cplusplus )
}
#endif

#endif /* SIXTRACKLIB_AMD_BEAM_ELEMENTS_H__ */

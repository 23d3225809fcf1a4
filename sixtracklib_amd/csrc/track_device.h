// Element maps on register-resident particle state (device code).
//
// One lane tracks one particle; the phase-space coordinates live in the
// ``Lane`` struct (registers) for the whole launch and touch HBM only at
// load, store and output time.  Element parameters come from the packed
// lattice through a constant-address-space pointer: every lane reads the
// same record, so the loads are scalar (SGPR) loads and the type dispatch
// is a scalar branch.
//
// Each map restates the reference map cited above it with the IEEE
// operations in the same order, so a build without floating-point
// contraction reproduces the reference bit for bit (transcendentals from
// include/sixtracklib_amd/stmath.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sixtracklib_amd/stmath.h"
#include "lattice.h"

#define ST_DEV __device__ __forceinline__
#define ST_DEV_NOINLINE __device__ __noinline__

typedef __attribute__( ( address_space( 4 ) ) ) const uint64_t cword_t;
typedef cword_t* cptr_t;

namespace st
{
constexpr double C_LIGHT = 299792458.0;
constexpr double EPSILON_0 = 8.854187817620e-12;
constexpr double PI = 3.1415926535897932384626433832795028841971693993751;
constexpr double SQRT_PI = 1.7724538509055160272981674833411451827975494561224;
constexpr double DEG2RAD_CONST = 0.0174532925199432957692369076848861271344287188854;
constexpr double QELEM = 1.60217662e-19;
constexpr double CHARGE0_SI = 1.602176634e-19;

// Lane bookkeeping lives in wave-uniform 64-bit masks (SGPRs): the vote of
// a predicate and the test of this lane's bit in a mask.
ST_DEV uint64_t ballot( bool p ) { return __builtin_amdgcn_ballot_w64( p ); }
ST_DEV bool lane_in( uint64_t m ) { return ( ( m >> __lane_id() ) & 1u ) != 0u; }

ST_DEV double rd( cptr_t p, int k ) { return __longlong_as_double( ( long long )p[ k ] ); }
ST_DEV int64_t ri( cptr_t p, int k ) { return ( int64_t )p[ k ]; }

struct Lane
{
    double x, px, y, py, zeta, delta, rpp, rvv, psigma, s;
    int64_t state;
};

struct LaneConst
{
    double q0, mass0, beta0, p0c, chi, charge_ratio;
};

// be_drift/track.h:45-67 without the path length s (drift() adds it).
// dzeta = rvv - (1 + (xp^2 + yp^2)/2) cancels: 1 + t/2 rounds t/2 (about
// 1e-10) to the spacing of doubles near 1, so the zeta gained per drift
// carries a relative rounding error of order 1e-7.  The FMA build keeps
// that rounding (1 + t/2 rounded once, as fma(t, 0.5, 1) does since t * 0.5
// is exact) and fuses the updates of x, y and zeta, which move each result
// by at most an ulp of its own size; t = xp^2 + yp^2 is one fused
// multiply-add there, an ulp of t (~1e-26) from the reference's, which
// changes 1 + t/2 only when that sum lies within 1e-26 of a rounding
// boundary.
ST_DEV void drift_xz( Lane& p, double const length )
{
    double const xp = p.px * p.rpp;
    double const yp = p.py * p.rpp;
#if defined( ST_FMA_BUILD )
    double const t = fma( xp, xp, yp * yp );
    double const dzeta = p.rvv - fma( t, 0.5, 1.0 );
    p.x = fma( xp, length, p.x );
    p.y = fma( yp, length, p.y );
    p.zeta = fma( length, dzeta, p.zeta );
#else
    double const dzeta = p.rvv - ( 1.0 + ( xp * xp + yp * yp ) / 2.0 );
    p.x += xp * length;
    p.y += yp * length;
    p.zeta += length * dzeta;
#endif
}

// be_drift/track.h:45-67
ST_DEV void drift( Lane& p, double const length )
{
    drift_xz( p, length );
    p.s += length;
}

// be_drift/track.h:71-97
ST_DEV void drift_exact( Lane& p, double const length )
{
#pragma clang fp contract( off )   // rvv * length - opd * lpzi cancels
    double const opd = p.delta + 1.0;
    double const lpzi = length / sqrt( opd * opd - ( p.px * p.px + p.py * p.py ) );
    double const dzeta = p.rvv * length - opd * lpzi;
    p.x += p.px * lpzi;
    p.y += p.py * lpzi;
    p.s += length;
    p.zeta += dzeta;
}

// be_limit/track.h:54-88: the particle survives iff sign(x)*x < 1 and
// sign(y)*y < 1, which is |x| < 1 and |y| < 1 for every input (+-0, +-Inf
// and NaN included).  Returns the survival predicate; the caller applies
// state &= ok.
ST_DEV bool limit_global_ok( const Lane& p )
{
    return ( fabs( p.x ) < 1.0 ) & ( fabs( p.y ) < 1.0 );
}

// The same test as a wave mask of the lanes that fail it: one compare per
// coordinate, each written straight into an SGPR pair.
ST_DEV uint64_t limit_global_lost_mask( const Lane& p )
{
    return __builtin_amdgcn_ballot_w64( !( fabs( p.x ) < 1.0 ) ) |
           __builtin_amdgcn_ballot_w64( !( fabs( p.y ) < 1.0 ) );
}

// One Horner step of be_multipole/track.h:62-76 (dpx + i dpy) <- bal +
// (dpx + i dpy)(x + i y).  The exact build keeps the reference's order of
// operations; the FMA build folds each component into two fused
// multiply-adds: fma(dpx, x, fma(-dpy, y, bal_re)).
#if defined( ST_FMA_BUILD )
// a * b + c and -a * b + c with c a wave-uniform (SGPR) value.  Written as
// one VOP3 v_fma_f64 reading c from its SGPR pair; left to itself the
// compiler selects the two-address v_fmac_f64 and first copies c into the
// destination VGPRs (two extra v_mov per call).
ST_DEV double fma_sc( double const a, double const b, double const c )
{
#if defined( __HIP_DEVICE_COMPILE__ )
    double d;
    asm( "v_fma_f64 %0, %1, %2, %3" : "=v"( d ) : "v"( a ), "v"( b ), "s"( c ) );
    return d;
#else
    return fma( a, b, c );
#endif
}

ST_DEV double fnma_sc( double const a, double const b, double const c )
{
#if defined( __HIP_DEVICE_COMPILE__ )
    double d;
    asm( "v_fma_f64 %0, -%1, %2, %3" : "=v"( d ) : "v"( a ), "v"( b ), "s"( c ) );
    return d;
#else
    return fma( -a, b, c );
#endif
}
#endif

ST_DEV void horner_step( double& dpx, double& dpy, double const x,
                         double const y, double const bre, double const bim )
{
#if defined( ST_FMA_BUILD )
    double const ndpx = fma( dpx, x, fnma_sc( dpy, y, bre ) );
    dpy = fma( dpx, y, fma_sc( dpy, x, bim ) );
    dpx = ndpx;
#else
    double const zre = dpx * x - dpy * y;
    double const zim = dpx * y + dpy * x;
    dpx = bre + zre;
    dpy = bim + zim;
#endif
}

// Horner over a compile-time order: all 2N+2 coefficients are read with
// constant offsets, so they arrive in one batch of scalar loads.
template< int N >
ST_DEV void horner_fixed( cptr_t e, double const x, double const y,
                          double& dpx, double& dpy )
{
    dpx = rd( e, 2 * N );
    dpy = rd( e, 2 * N + 1 );
#pragma unroll
    for( int k = N; k > 0; --k )
        horner_step( dpx, dpy, x, y, rd( e, 2 * k - 2 ), rd( e, 2 * k - 1 ) );
}

ST_DEV void horner( cptr_t e, int const order, double const x, double const y,
                    double& dpx, double& dpy )
{
    switch( order )
    {
    case 0: horner_fixed< 0 >( e, x, y, dpx, dpy ); return;
    case 1: horner_fixed< 1 >( e, x, y, dpx, dpy ); return;
    case 2: horner_fixed< 2 >( e, x, y, dpx, dpy ); return;
    case 3: horner_fixed< 3 >( e, x, y, dpx, dpy ); return;
    case 4: horner_fixed< 4 >( e, x, y, dpx, dpy ); return;
    case 5: horner_fixed< 5 >( e, x, y, dpx, dpy ); return;
    case 6: horner_fixed< 6 >( e, x, y, dpx, dpy ); return;
    case 7: horner_fixed< 7 >( e, x, y, dpx, dpy ); return;
    case 8: horner_fixed< 8 >( e, x, y, dpx, dpy ); return;
    case 9: horner_fixed< 9 >( e, x, y, dpx, dpy ); return;
    case 10: horner_fixed< 10 >( e, x, y, dpx, dpy ); return;
    default: break;
    }
    int ix = 2 * order;
    dpx = rd( e, ix );
    dpy = rd( e, ix + 1 );
    for( ; ix > 0; ix -= 2 ) horner_step( dpx, dpy, x, y, rd( e, ix - 2 ), rd( e, ix - 1 ) );
}

// be_multipole/track.h:38-112
ST_DEV void multipole( Lane& p, double const chi, cptr_t e, unsigned aux,
                       double const delta )
{
    int const order = ( int )( aux & 0x7fffu );
    double dpx, dpy;
    double const x = p.x, y = p.y;
    horner( e, order, x, y, dpx, dpy );
#if defined( ST_FMA_BUILD )
    // px -= chi dpx, py += chi dpy as one fused step each (the common case;
    // contraction does not reach across the curvature branch by itself)
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
        double const hxl = rd( e, nb ), hyl = rd( e, nb + 1 ), length = rd( e, nb + 2 );
        double const hxlx = x * hxl;
        double const hyly = y * hyl;
        p.zeta += chi * ( hyly - hxlx );
        dpx += hxl + hxl * delta;
        dpy -= hyl + hyl * delta;
        if( length > 0.0 )
        {
            double const b1l = chi * rd( e, 0 );
            double const a1l = chi * rd( e, 1 );
            dpx -= b1l * hxlx / length;
            dpy += a1l * hyly / length;
        }
    }
    p.px += dpx;
    p.py += dpy;
}

ST_DEV void multipole( Lane& p, const LaneConst& c, cptr_t e, unsigned aux,
                       double const delta )
{
    multipole( p, c.chi, e, aux, delta );
}

// particles.h:5255-5297 (energy0 5220-5234, zeta scaling 5024-5032)
ST_DEV void add_to_energy( Lane& p, const LaneConst& c, double const dE )
{
#pragma clang fp contract( off )
    double const beta0 = c.beta0;
    double const delta_beta0 = p.delta * beta0;
    double const energy0 = sqrt( c.p0c * c.p0c + c.mass0 * c.mass0 );
    double const ptau_beta0 = dE / energy0 +
        sqrt( delta_beta0 * delta_beta0 + 2.0 * delta_beta0 * beta0 + 1.0 ) - 1.0;
    double const ptau = ptau_beta0 / beta0;
    double const psigma = ptau / beta0;
    double const delta = sqrt( ptau * ptau + 2.0 * psigma + 1.0 ) - 1.0;
    double const one_plus_delta = delta + 1.0;
    double const rvv = one_plus_delta / ( 1.0 + ptau_beta0 );
    p.delta = delta;
    p.psigma = psigma;
    p.zeta *= rvv / p.rvv;
    p.rvv = rvv;
    p.rpp = 1.0 / one_plus_delta;
}

// particles.h:5459-5497
ST_DEV void update_delta( Lane& p, const LaneConst& c, double const new_delta )
{
#pragma clang fp contract( off )
    double const beta0 = c.beta0;
    double const delta_beta0 = new_delta * beta0;
    double const ptau_beta0 = sqrt( delta_beta0 * delta_beta0 +
        2.0 * delta_beta0 * beta0 + 1.0 ) - 1.0;
    double const one_plus_delta = 1.0 + new_delta;
    p.delta = new_delta;
    p.rvv = one_plus_delta / ( 1.0 + ptau_beta0 );
    p.rpp = 1.0 / one_plus_delta;
    p.psigma = ptau_beta0 / ( beta0 * beta0 );
}

// be_cavity/track.h:39-67 with DEG2RAD*lag and K_FACTOR*f folded at pack time
ST_DEV void cavity( Lane& p, const LaneConst& c, cptr_t e )
{
#pragma clang fp contract( off )
    double const voltage = rd( e, 0 ), lag_rad = rd( e, 1 ), kfreq = rd( e, 2 );
    double const tau = p.zeta / ( c.beta0 * p.rvv );
    double const phase = lag_rad - kfreq * tau;
    double const q = c.q0 * c.charge_ratio;
    double const energy = q * voltage * st_sin( phase );
    add_to_energy( p, c, energy );
}

// be_xyshift/track.h:38-55 (payload holds -dx, -dy)
ST_DEV void xyshift( Lane& p, cptr_t e )
{
    p.x += rd( e, 0 );
    p.y += rd( e, 1 );
}

// be_srotation/track.h:38-69
ST_DEV void srotation( Lane& p, cptr_t e )
{
    double const cos_z = rd( e, 0 ), sin_z = rd( e, 1 );
    double const x = p.x, y = p.y, px = p.px, py = p.py;
    p.x = cos_z * x + sin_z * y;
    p.y = -sin_z * x + cos_z * y;
    p.px = cos_z * px + sin_z * py;
    p.py = -sin_z * px + cos_z * py;
}

// Aperture predicates (the caller applies state &= ok,
// particles.h:6493-6505).  The two ellipse tests switch contraction off in
// the FMA build as well: a fused x^2 b^2 + y^2 a^2 rounds differently from
// the reference's products and sum, and a particle on the boundary could
// then be kept where the reference loses it (or lost where it is kept).
// be_limit/track.h:90-111
ST_DEV bool limit_rect_ok( const Lane& p, cptr_t e )
{
    return ( p.x >= rd( e, 0 ) ) && ( p.x <= rd( e, 1 ) ) &&
           ( p.y >= rd( e, 2 ) ) && ( p.y <= rd( e, 3 ) );
}

// be_limit/track.h:113-137
ST_DEV bool limit_ellipse_ok( const Lane& p, cptr_t e )
{
#pragma clang fp contract( off )
    double temp = p.x;
    double y_squ = p.y;
    temp *= temp;
    temp *= rd( e, 1 );
    y_squ *= y_squ;
    y_squ *= rd( e, 0 );
    temp += y_squ;
    return temp <= rd( e, 2 );
}

// be_limit/track.h:139-162
ST_DEV bool limit_rect_ellipse_ok( const Lane& p, cptr_t e )
{
#pragma clang fp contract( off )
    double const x = p.x, y = p.y;
    double const max_x = rd( e, 0 ), max_y = rd( e, 1 );
    double temp = x * x * rd( e, 3 );
    temp += y * y * rd( e, 2 );
    return ( x <= max_x ) && ( x >= -max_x ) && ( y <= max_y ) &&
           ( y >= -max_y ) && ( temp <= rd( e, 4 ) );
}

// be_dipedge/track.h:37-51
ST_DEV void dipedge( Lane& p, cptr_t e )
{
    p.px += rd( e, 0 ) * p.x;
    p.py += rd( e, 1 ) * p.y;
}

// From here to the space-charge maps every function switches contraction
// off, also in the FMA build: the field formulas cancel (1 - exp(-r^2/2s^2)
// at small r, differences of the beam-size terms), so a fused multiply-add
// there changes a kick by far more than an ulp of the reference's.  These
// maps run a few times per turn and cost nothing measurable.
// be_rfmultipole/track.h:39-118
ST_DEV_NOINLINE void rf_multipole( Lane& p, const LaneConst& c, cptr_t e,
                                   unsigned aux )
{
#pragma clang fp contract( off )
    int const order = ( int )aux;
    int const nb = 2 * order + 2;
    double const voltage = rd( e, 0 ), frequency = rd( e, 1 ), lag = rd( e, 2 );
    double const DEG2RAD = PI / 180.0;
    double const k = frequency * ( 2.0 * PI ) / C_LIGHT;
    double const x = p.x, y = p.y;
    double const q = c.q0 * c.charge_ratio * 1.0;
    double const ktau = k * p.zeta / ( c.beta0 * p.rvv );
    double dpx = 0.0, dpy = 0.0, dptr = 0.0, zre = 1.0, zim = 0.0;
    for( int kk = 0; kk <= order; kk++ )
    {
        double const pn_kk = DEG2RAD_CONST * rd( e, 3 + nb + 2 * kk ) - ktau;
        double const ps_kk = DEG2RAD * rd( e, 3 + nb + 2 * kk + 1 ) - ktau;
        double const bal_n_kk = rd( e, 3 + 2 * kk );
        double const bal_s_kk = rd( e, 3 + 2 * kk + 1 );
        double const cn = st_cos( pn_kk );
        double const cs = st_cos( ps_kk );
        double const sn = st_sin( pn_kk );
        double const ss = st_sin( ps_kk );
        dpx += cn * ( bal_n_kk * zre ) - cs * ( bal_s_kk * zim );
        dpy += cs * ( bal_s_kk * zre ) + cn * ( bal_n_kk * zim );
        double const zret = zre * x - zim * y;
        zim = zim * x + zre * y;
        zre = zret;
        dptr += sn * ( bal_n_kk * zre ) - ss * ( bal_s_kk * zim );
    }
    double const cav_energy = q * ( voltage * st_sin( lag * DEG2RAD_CONST - ktau ) );
    double const rfmultipole_energy = -q * ( ( k * c.p0c ) * dptr );
    double const px_kick = -c.chi * dpx;
    double const py_kick = c.chi * dpy;
    double const energy_kick = cav_energy + rfmultipole_energy;
    p.px += px_kick;
    p.py += py_kick;
    add_to_energy( p, c, energy_kick );
}

// ---------------------------------------------------------------- fields

// CERN C335 complex error function, be_beamfields/faddeeva_cern.h:45-120.
// The reference fills Rx/Ry[0..nu] in a first descending loop and reads
// Rx/Ry[nc-1..0] in a second descending loop; since nc < nu the second
// loop's step n can run right after the first loop produced R[n-1], which
// removes the per-lane arrays (scratch memory) with the identical sequence
// of IEEE operations.
ST_DEV void cerrf( double const in_real, double const in_imag, double& out_re,
                   double& out_im )
{
#pragma clang fp contract( off )
    double const a_constant = 1.12837916709551;
    double const xLim = 5.33;
    double const yLim = 4.29;
    double Wx, Wy;
    double const x = fabs( in_real );
    double const y = fabs( in_imag );
    if( y < yLim && x < xLim )
    {
        double const q = ( 1.0 - y / yLim ) * sqrt( 1.0 - ( x / xLim ) * ( x / xLim ) );
        double const h = 1.0 / ( 3.2 * q );
        int const nc = 7 + ( int )( 23.0 * q );
        double xl = st_pow_int( h, 1 - nc );
        double const xh = y + 0.5 / h;
        double const yh = x;
        int const nu = 10 + ( int )( 21.0 * q );
        double Rx = 0.0, Ry = 0.0, Sx = 0.0, Sy = 0.0;
        for( int n = nu; n > 0; n-- )
        {
            double const Tx = xh + n * Rx;
            double const Ty = yh - n * Ry;
            double const Tn = Tx * Tx + Ty * Ty;
            Rx = 0.5 * Tx / Tn;                 // = R[n-1]
            Ry = 0.5 * Ty / Tn;
            if( n <= nc )
            {
                double const Saux = Sx + xl;
                double const Sx_new = Rx * Saux - Ry * Sy;
                Sy = Rx * Sy + Ry * Saux;
                Sx = Sx_new;
                xl = h * xl;
            }
        }
        Wx = a_constant * Sx;
        Wy = a_constant * Sy;
    }
    else
    {
        double const xh = y;
        double const yh = x;
        double Rx = 0.0, Ry = 0.0;
        for( int n = 9; n > 0; n-- )
        {
            double const Tx = xh + n * Rx;
            double const Ty = yh - n * Ry;
            double const Tn = Tx * Tx + Ty * Ty;
            Rx = 0.5 * Tx / Tn;
            Ry = 0.5 * Ty / Tn;
        }
        Wx = a_constant * Rx;
        Wy = a_constant * Ry;
    }
    if( y == 0. ) { Wx = st_exp( -x * x ); }
    if( in_imag < 0. )
    {
        Wx =   2.0 * st_exp( y * y - x * x ) * st_cos( 2.0 * x * y ) - Wx;
        Wy = - 2.0 * st_exp( y * y - x * x ) * st_sin( 2.0 * x * y ) - Wy;
        if( in_real > 0. ) { Wy = -Wy; }
    }
    else if( in_real < 0. ) { Wy = -Wy; }
    out_re = Wx;
    out_im = Wy;
}

// be_beamfields/gauss_fields.h:56-70
ST_DEV void field_round( double const sigma, double const x, double const y,
                         double& Ex, double& Ey )
{
#pragma clang fp contract( off )
    double const Delta_x = 0., Delta_y = 0.;
    double const r2 = ( x - Delta_x ) * ( x - Delta_x ) + ( y - Delta_y ) * ( y - Delta_y );
    double temp;
    if( r2 < 1e-20 ) temp = sqrt( r2 ) / ( 2. * PI * EPSILON_0 * sigma );
    else temp = ( 1 - st_exp( -0.5 * r2 / ( sigma * sigma ) ) ) /
                ( 2. * PI * EPSILON_0 * r2 );
    Ex = temp * ( x - Delta_x );
    Ey = temp * ( y - Delta_y );
}

// be_beamfields/gauss_fields.h:72-146
ST_DEV void field_ellip( double const sigmax, double const sigmay,
                         double const x, double const y, double& Ex_out,
                         double& Ey_out )
{
#pragma clang fp contract( off )
    double const abx = fabs( x - 0. );
    double const aby = fabs( y - 0. );
    double Ex, Ey;
    if( sigmax > sigmay )
    {
        double const S = sqrt( 2. * ( sigmax * sigmax - sigmay * sigmay ) );
        double const factBE = 1. / ( 2. * EPSILON_0 * SQRT_PI * S );
        double const etaBE_re = sigmay / sigmax * abx;
        double const etaBE_im = sigmax / sigmay * aby;
        double wz_re, wz_im, we_re, we_im;
        cerrf( abx / S, aby / S, wz_re, wz_im );
        cerrf( etaBE_re / S, etaBE_im / S, we_re, we_im );
        double const expBE = st_exp( -abx * abx / ( 2 * sigmax * sigmax ) -
                                     aby * aby / ( 2 * sigmay * sigmay ) );
        Ex = factBE * ( wz_im - we_im * expBE );
        Ey = factBE * ( wz_re - we_re * expBE );
    }
    else if( sigmax < sigmay )
    {
        double const S = sqrt( 2. * ( sigmay * sigmay - sigmax * sigmax ) );
        double const factBE = 1. / ( 2. * EPSILON_0 * SQRT_PI * S );
        double const etaBE_re = sigmax / sigmay * aby;
        double const etaBE_im = sigmay / sigmax * abx;
        double wz_re, wz_im, we_re, we_im;
        cerrf( aby / S, abx / S, wz_re, wz_im );
        cerrf( etaBE_re / S, etaBE_im / S, we_re, we_im );
        double const expBE = st_exp( -aby * aby / ( 2 * sigmay * sigmay ) -
                                     abx * abx / ( 2 * sigmax * sigmax ) );
        Ey = factBE * ( wz_im - we_im * expBE );
        Ex = factBE * ( wz_re - we_re * expBE );
    }
    else
    {
        Ex = Ey = __builtin_inf();
    }
    if( ( x - 0. ) < 0 ) Ex = -Ex;
    if( ( y - 0. ) < 0 ) Ey = -Ey;
    Ex_out = Ex;
    Ey_out = Ey;
}

// be_beamfields/gauss_fields.h:150-208
template< bool SKIP_GS >
ST_DEV void gauss_field( double const x, double const y, double const sigma_x,
                         double const sigma_y, double const min_sigma_diff,
                         double& Ex, double& Ey, double& Gx, double& Gy )
{
#pragma clang fp contract( off )
    if( fabs( sigma_x - sigma_y ) < min_sigma_diff )
    {
        double const sigma = 0.5 * ( sigma_x + sigma_y );
        field_round( sigma, x, y, Ex, Ey );
        if( SKIP_GS ) { Gx = 0.; Gy = 0.; }
        else
        {
            Gx = 1 / ( 2. * ( x * x + y * y ) ) * ( y * Ey - x * Ex +
                1. / ( 2 * PI * EPSILON_0 * sigma * sigma ) * x * x *
                st_exp( -( x * x + y * y ) / ( 2. * sigma * sigma ) ) );
            Gy = 1. / ( 2 * ( x * x + y * y ) ) * ( x * Ex - y * Ey +
                1. / ( 2 * PI * EPSILON_0 * sigma * sigma ) * y * y *
                st_exp( -( x * x + y * y ) / ( 2. * sigma * sigma ) ) );
        }
    }
    else
    {
        field_ellip( sigma_x, sigma_y, x, y, Ex, Ey );
        double const Sig_11 = sigma_x * sigma_x;
        double const Sig_33 = sigma_y * sigma_y;
        if( SKIP_GS ) { Gx = 0.; Gy = 0.; }
        else
        {
            Gx = -1. / ( 2 * ( Sig_11 - Sig_33 ) ) * ( x * Ex + y * Ey +
                1. / ( 2 * PI * EPSILON_0 ) * ( sigma_y / sigma_x *
                st_exp( -x * x / ( 2 * Sig_11 ) - y * y / ( 2 * Sig_33 ) ) - 1. ) );
            Gy = 1. / ( 2 * ( Sig_11 - Sig_33 ) ) * ( x * Ex + y * Ey +
                1. / ( 2 * PI * EPSILON_0 ) * ( sigma_x / sigma_y *
                st_exp( -x * x / ( 2 * Sig_11 ) - y * y / ( 2 * Sig_33 ) ) - 1. ) );
        }
    }
}

// be_beamfields/track.h:89-149 (data words: q_part, N_part, sigma_x,
// sigma_y, beta_s, min_sigma_diff, Delta_x, Delta_y, Dpx_sub, Dpy_sub, enabled)
ST_DEV_NOINLINE void bb4d( Lane& p, const LaneConst& c, cptr_t d )
{
#pragma clang fp contract( off )
    if( !( rd( d, 10 ) > 0. ) ) return;
    double const charge = 1. * c.q0 * QELEM;
    double const x = p.x - rd( d, 6 );
    double const y = p.y - rd( d, 7 );
    double const beta = c.beta0 / p.rvv;
    double const p0c = c.p0c * QELEM;
    double Ex, Ey, Gx, Gy;
    gauss_field< true >( x, y, rd( d, 2 ), rd( d, 3 ), rd( d, 5 ), Ex, Ey, Gx, Gy );
    double const beta_s = rd( d, 4 );
    double const fact_kick = c.chi * rd( d, 1 ) * rd( d, 0 ) * charge *
        ( 1. + beta * beta_s ) / ( p0c * ( beta + beta_s ) );
    p.px += ( fact_kick * Ex - rd( d, 8 ) );
    p.py += ( fact_kick * Ey - rd( d, 9 ) );
}

ST_DEV double sgn( double a ) { return ( double )( ( ( a ) >= 0 ) - ( ( a ) < 0 ) ); }

// be_beamfields.h:2237-2385 (data: Sig_11_0 .. Sig_44_0 at words 6..15)
ST_DEV void bb6d_sigma( cptr_t d, double const S, double const threshold_singular,
    double& Sig_11_hat, double& Sig_33_hat, double& costheta, double& sintheta,
    double& dS_Sig_11_hat, double& dS_Sig_33_hat, double& dS_costheta,
    double& dS_sintheta )
{
#pragma clang fp contract( off )
    double const Sig_11_0 = rd( d, 6 ), Sig_12_0 = rd( d, 7 ), Sig_13_0 = rd( d, 8 );
    double const Sig_14_0 = rd( d, 9 ), Sig_22_0 = rd( d, 10 ), Sig_23_0 = rd( d, 11 );
    double const Sig_24_0 = rd( d, 12 ), Sig_33_0 = rd( d, 13 ), Sig_34_0 = rd( d, 14 );
    double const Sig_44_0 = rd( d, 15 );
    double const Sig_11 = Sig_11_0 + 2. * Sig_12_0 * S + Sig_22_0 * S * S;
    double const Sig_33 = Sig_33_0 + 2. * Sig_34_0 * S + Sig_44_0 * S * S;
    double const Sig_13 = Sig_13_0 + ( Sig_14_0 + Sig_23_0 ) * S + Sig_24_0 * S * S;
    double const Sig_12 = Sig_12_0 + Sig_22_0 * S;
    double const Sig_14 = Sig_14_0 + Sig_24_0 * S;
    double const Sig_22 = Sig_22_0 + 0. * S;
    double const Sig_23 = Sig_23_0 + Sig_24_0 * S;
    double const Sig_24 = Sig_24_0 + 0. * S;
    double const Sig_34 = Sig_34_0 + Sig_44_0 * S;
    double const Sig_44 = Sig_44_0 + 0. * S;
    double const R = Sig_11 - Sig_33;
    double const W = Sig_11 + Sig_33;
    double const T = R * R + 4 * Sig_13 * Sig_13;
    double const dS_R = 2. * ( Sig_12_0 - Sig_34_0 ) + 2 * S * ( Sig_22_0 - Sig_44_0 );
    double const dS_W = 2. * ( Sig_12_0 + Sig_34_0 ) + 2 * S * ( Sig_22_0 + Sig_44_0 );
    double const dS_Sig_13 = Sig_14_0 + Sig_23_0 + 2 * Sig_24_0 * S;
    double const dS_T = 2 * R * dS_R + 8. * Sig_13 * dS_Sig_13;
    double const signR = sgn( R );
    double cos2theta, dS_cos2theta;
    if( T < threshold_singular )
    {
        double const a = Sig_12 - Sig_34;
        double const b = Sig_22 - Sig_44;
        double const cc = Sig_14 + Sig_23;
        double const dd = Sig_24;
        double sqrt_a2_c2 = sqrt( a * a + cc * cc );
        if( sqrt_a2_c2 * sqrt_a2_c2 * sqrt_a2_c2 < threshold_singular )
        {
            if( fabs( dd ) > threshold_singular )
                cos2theta = fabs( b ) / sqrt( b * b + 4 * dd * dd );
            else
                cos2theta = 1.;
            costheta = sqrt( 0.5 * ( 1. + cos2theta ) );
            sintheta = sgn( b ) * sgn( dd ) * sqrt( 0.5 * ( 1. - cos2theta ) );
            dS_costheta = 0.;
            dS_sintheta = 0.;
            Sig_11_hat = 0.5 * W;
            Sig_33_hat = 0.5 * W;
            dS_Sig_11_hat = 0.5 * dS_W;
            dS_Sig_33_hat = 0.5 * dS_W;
        }
        else
        {
            sqrt_a2_c2 = sqrt( a * a + cc * cc );
            cos2theta = fabs( 2. * a ) / ( 2 * sqrt_a2_c2 );
            costheta = sqrt( 0.5 * ( 1. + cos2theta ) );
            sintheta = sgn( a ) * sgn( cc ) * sqrt( 0.5 * ( 1. - cos2theta ) );
            dS_cos2theta = sgn( a ) * ( 0.5 * b / sqrt_a2_c2 - a * ( a * b +
                2. * cc * dd ) / ( 2. * sqrt_a2_c2 * sqrt_a2_c2 * sqrt_a2_c2 ) );
            dS_costheta = 1. / ( 4. * costheta ) * dS_cos2theta;
            if( fabs( sintheta ) > threshold_singular )
                dS_sintheta = -1. / ( 4. * sintheta ) * dS_cos2theta;
            else
                dS_sintheta = dd / ( 2. * a );
            Sig_11_hat = 0.5 * W;
            Sig_33_hat = 0.5 * W;
            dS_Sig_11_hat = 0.5 * dS_W + sgn( a ) * sqrt_a2_c2;
            dS_Sig_33_hat = 0.5 * dS_W - sgn( a ) * sqrt_a2_c2;
        }
    }
    else
    {
        double const sqrtT = sqrt( T );
        cos2theta = signR * R / sqrtT;
        costheta = sqrt( 0.5 * ( 1. + cos2theta ) );
        sintheta = signR * sgn( Sig_13 ) * sqrt( 0.5 * ( 1. - cos2theta ) );
        Sig_11_hat = 0.5 * ( W + signR * sqrtT );
        Sig_33_hat = 0.5 * ( W - signR * sqrtT );
        dS_cos2theta = signR * ( dS_R / sqrtT - R / ( 2 * sqrtT * sqrtT * sqrtT ) * dS_T );
        dS_costheta = 1. / ( 4. * costheta ) * dS_cos2theta;
        if( fabs( sintheta ) < threshold_singular )
            dS_sintheta = ( Sig_14 + Sig_23 ) / R;
        else
            dS_sintheta = -1. / ( 4. * sintheta ) * dS_cos2theta;
        dS_Sig_11_hat = 0.5 * ( dS_W + signR * 0.5 / sqrtT * dS_T );
        dS_Sig_33_hat = 0.5 * ( dS_W - signR * 0.5 / sqrtT * dS_T );
    }
}

// be_beamfields/track.h:342-507 with boost be_beamfields.h:2101-2161 and
// inverse boost 2163-2235.  Data word map (NS(BB6D_data)): 0 q_part,
// 1-5 sphi cphi tphi salpha calpha, 6-15 sigmas, 16 min_sigma_diff,
// 17 threshold_singular, 18 N_slices, 19-20 delta_x/y, 21-26 CO,
// 27-32 D*_sub, 33 enabled, 38.. slices (N, x*, y*, sigma*).
ST_DEV_NOINLINE void bb6d( Lane& p, const LaneConst& c, cptr_t d )
{
#pragma clang fp contract( off )
    if( !( rd( d, 33 ) > 0. ) ) return;
    int const N_slices = ( int )ri( d, 18 );
    cptr_t N_part_per_slice = d + 38;
    cptr_t x_slices_star = d + 38 + N_slices;
    cptr_t y_slices_star = d + 38 + 2 * N_slices;
    cptr_t sigma_slices_star = d + 38 + 3 * N_slices;
    double const sphi = rd( d, 1 ), cphi = rd( d, 2 ), tphi = rd( d, 3 );
    double const salpha = rd( d, 4 ), calpha = rd( d, 5 );
    double const q0 = QELEM * c.q0;
    double const P0 = c.p0c / C_LIGHT * QELEM;
    double x_star = p.x - rd( d, 21 ) - rd( d, 19 );
    double px_star = p.px - rd( d, 22 );
    double y_star = p.y - rd( d, 23 ) - rd( d, 20 );
    double py_star = p.py - rd( d, 24 );
    double sigma_star = p.zeta - rd( d, 25 );
    double delta_star = p.delta - rd( d, 26 );
    {   // boost
        double const x = x_star, px = px_star, y = y_star, py = py_star;
        double const sigma = sigma_star, delta = delta_star;
        double const h = delta + 1. - sqrt( ( 1. + delta ) * ( 1. + delta ) - px * px - py * py );
        double const px_st = px / cphi - h * calpha * tphi / cphi;
        double const py_st = py / cphi - h * salpha * tphi / cphi;
        double const delta_st = delta - px * calpha * tphi - py * salpha * tphi + h * tphi * tphi;
        double const pz_st = sqrt( ( 1. + delta_st ) * ( 1. + delta_st ) - px_st * px_st - py_st * py_st );
        double const hx_st = px_st / pz_st;
        double const hy_st = py_st / pz_st;
        double const hsigma_st = 1. - ( delta_st + 1 ) / pz_st;
        double const L11 = 1. + hx_st * calpha * sphi;
        double const L12 = hx_st * salpha * sphi;
        double const L13 = calpha * tphi;
        double const L21 = hy_st * calpha * sphi;
        double const L22 = 1. + hy_st * salpha * sphi;
        double const L23 = salpha * tphi;
        double const L31 = hsigma_st * calpha * sphi;
        double const L32 = hsigma_st * salpha * sphi;
        double const L33 = 1. / cphi;
        x_star = L11 * x + L12 * y + L13 * sigma;
        y_star = L21 * x + L22 * y + L23 * sigma;
        sigma_star = L31 * x + L32 * y + L33 * sigma;
        px_star = px_st;
        py_star = py_st;
        delta_star = delta_st;
    }
    double const q_part = rd( d, 0 );
    double const threshold_singular = rd( d, 17 );
    double const min_sigma_diff = rd( d, 16 );
    for( int i_slice = 0; i_slice < N_slices; i_slice++ )
    {
        double const sigma_slice_star = rd( sigma_slices_star, i_slice );
        double const x_slice_star = rd( x_slices_star, i_slice );
        double const y_slice_star = rd( y_slices_star, i_slice );
        double const Ksl = rd( N_part_per_slice, i_slice ) * q_part * q0 / ( P0 * C_LIGHT );
        double const S = 0.5 * ( sigma_star - sigma_slice_star );
        double Sig_11_hat_star, Sig_33_hat_star, costheta, sintheta;
        double dS_Sig_11_hat_star, dS_Sig_33_hat_star, dS_costheta, dS_sintheta;
        bb6d_sigma( d, S, threshold_singular, Sig_11_hat_star, Sig_33_hat_star,
                    costheta, sintheta, dS_Sig_11_hat_star, dS_Sig_33_hat_star,
                    dS_costheta, dS_sintheta );
        double const x_bar_star = x_star + px_star * S - x_slice_star;
        double const y_bar_star = y_star + py_star * S - y_slice_star;
        double const x_bar_hat_star = x_bar_star * costheta + y_bar_star * sintheta;
        double const y_bar_hat_star = -x_bar_star * sintheta + y_bar_star * costheta;
        double const dS_x_bar_hat_star = x_bar_star * dS_costheta + y_bar_star * dS_sintheta;
        double const dS_y_bar_hat_star = -x_bar_star * dS_sintheta + y_bar_star * dS_costheta;
        double Ex, Ey, Gx = 0.0, Gy = 0.0;
        gauss_field< false >( x_bar_hat_star, y_bar_hat_star, sqrt( Sig_11_hat_star ),
                              sqrt( Sig_33_hat_star ), min_sigma_diff, Ex, Ey, Gx, Gy );
        double const Fx_hat_star = Ksl * Ex;
        double const Fy_hat_star = Ksl * Ey;
        double const Gx_hat_star = Ksl * Gx;
        double const Gy_hat_star = Ksl * Gy;
        double const Fx_star = Fx_hat_star * costheta - Fy_hat_star * sintheta;
        double const Fy_star = Fx_hat_star * sintheta + Fy_hat_star * costheta;
        double const Fz_star = 0.5 * ( Fx_hat_star * dS_x_bar_hat_star +
                                       Fy_hat_star * dS_y_bar_hat_star +
                                       Gx_hat_star * dS_Sig_11_hat_star +
                                       Gy_hat_star * dS_Sig_33_hat_star );
        delta_star = delta_star + Fz_star + 0.5 * (
            Fx_star * ( px_star + 0.5 * Fx_star ) +
            Fy_star * ( py_star + 0.5 * Fy_star ) );
        x_star = x_star - S * Fx_star;
        px_star = px_star + Fx_star;
        y_star = y_star - S * Fy_star;
        py_star = py_star + Fy_star;
    }
    {   // inverse boost
        double const x_st = x_star, px_st = px_star, y_st = y_star, py_st = py_star;
        double const sigma_st = sigma_star, delta_st = delta_star;
        double const pz_st = sqrt( ( 1. + delta_st ) * ( 1. + delta_st ) - px_st * px_st - py_st * py_st );
        double const hx_st = px_st / pz_st;
        double const hy_st = py_st / pz_st;
        double const hsigma_st = 1. - ( delta_st + 1 ) / pz_st;
        double const Det_L = 1. / cphi + ( hx_st * calpha + hy_st * salpha - hsigma_st * sphi ) * tphi;
        double const Linv_11 = ( 1. / cphi + salpha * tphi * ( hy_st - hsigma_st * salpha * sphi ) ) / Det_L;
        double const Linv_12 = ( salpha * tphi * ( hsigma_st * calpha * sphi - hx_st ) ) / Det_L;
        double const Linv_13 = -tphi * ( calpha - hx_st * salpha * salpha * sphi +
                                         hy_st * calpha * salpha * sphi ) / Det_L;
        double const Linv_21 = ( calpha * tphi * ( -hy_st + hsigma_st * salpha * sphi ) ) / Det_L;
        double const Linv_22 = ( 1. / cphi + calpha * tphi * ( hx_st - hsigma_st * calpha * sphi ) ) / Det_L;
        double const Linv_23 = -tphi * ( salpha - hy_st * calpha * calpha * sphi +
                                         hx_st * calpha * salpha * sphi ) / Det_L;
        double const Linv_31 = -hsigma_st * calpha * sphi / Det_L;
        double const Linv_32 = -hsigma_st * salpha * sphi / Det_L;
        double const Linv_33 = ( 1. + hx_st * calpha * sphi + hy_st * salpha * sphi ) / Det_L;
        x_star = Linv_11 * x_st + Linv_12 * y_st + Linv_13 * sigma_st;
        y_star = Linv_21 * x_st + Linv_22 * y_st + Linv_23 * sigma_st;
        sigma_star = Linv_31 * x_st + Linv_32 * y_st + Linv_33 * sigma_st;
        double const h = ( delta_st + 1. - pz_st ) * cphi * cphi;
        px_star = px_st * cphi + h * calpha * tphi;
        py_star = py_st * cphi + h * salpha * tphi;
        delta_star = delta_st + px_star * calpha * tphi + py_star * salpha * tphi - h * tphi * tphi;
    }
    p.x = x_star + rd( d, 21 ) + rd( d, 19 ) - rd( d, 27 );
    p.px = px_star + rd( d, 22 ) - rd( d, 28 );
    p.y = y_star + rd( d, 23 ) + rd( d, 20 ) - rd( d, 29 );
    p.py = py_star + rd( d, 24 ) - rd( d, 30 );
    p.zeta = sigma_star + rd( d, 25 ) - rd( d, 31 );
    update_delta( p, c, delta_star + rd( d, 26 ) - rd( d, 32 ) );
}

// be_beamfields/track.h:154-198 (words: N, circumference, sigma_x,
// sigma_y, length, x_co, y_co, min_sigma_diff, enabled)
ST_DEV_NOINLINE void sc_coasting( Lane& p, const LaneConst& c, cptr_t e )
{
#pragma clang fp contract( off )
    double const charge = CHARGE0_SI * c.q0;
    double const beta0 = c.beta0;
    double const pc0 = c.p0c * CHARGE0_SI;
    double fact_kick = rd( e, 0 ) * rd( e, 4 ) * c.chi * c.charge_ratio *
        charge * charge * ( 1.0 - beta0 * beta0 );
    fact_kick /= rd( e, 1 ) * pc0 * beta0 * p.rvv;
    double Ex, Ey, Gx, Gy;
    gauss_field< true >( p.x - rd( e, 5 ), p.y - rd( e, 6 ), rd( e, 2 ), rd( e, 3 ),
                         rd( e, 7 ), Ex, Ey, Gx, Gy );
    p.px += fact_kick * Ex;
    p.py += fact_kick * Ey;
}

// be_beamfields/track.h:203-268 + internal/math_qgauss.h:114-179 (words:
// N, bunchlength_rms, sigma_x, sigma_y, length, x_co, y_co,
// min_sigma_diff, q_param, cq, enabled)
ST_DEV_NOINLINE void sc_qgauss( Lane& p, const LaneConst& c, cptr_t e )
{
#pragma clang fp contract( off )
    double const charge = CHARGE0_SI * c.q0;
    double const beta0 = c.beta0;
    double const beta = beta0 * p.rvv;
    double const pc0 = c.p0c * CHARGE0_SI;
    double const z = p.zeta / p.rvv;
    double const sqrt_beta = 1.0 / ( sqrt( 2.0 ) * rd( e, 1 ) );
    double fact_kick = c.chi * c.charge_ratio * rd( e, 4 ) * charge * charge *
        ( 1.0 - beta0 * beta );
    fact_kick *= rd( e, 0 );
    {
        double const q = rd( e, 8 ), cq = rd( e, 9 );
        double const arg = -( sqrt_beta * sqrt_beta * z * z );
        double const one_minus_q = 1.0 - q;
        double const abs_omq = ( one_minus_q >= 0.0 ) ? one_minus_q : -one_minus_q;
        double ev;
        if( abs_omq >= 1e-6 )
        {
            double u_plus = 1.0 + one_minus_q * arg;
            if( u_plus < 0.0 ) u_plus = 0.0;
            ev = st_pow( u_plus, 1.0 / one_minus_q );
        }
        else ev = st_exp( arg );
        fact_kick *= ( sqrt_beta / cq ) * ev;
    }
    fact_kick /= pc0 * beta;
    double Ex, Ey, Gx, Gy;
    gauss_field< true >( p.x - rd( e, 5 ), p.y - rd( e, 6 ), rd( e, 2 ), rd( e, 3 ),
                         rd( e, 7 ), Ex, Ey, Gx, Gy );
    p.px += fact_kick * Ex;
    p.py += fact_kick * Ey;
}

// be_beamfields/track.h:273-337 (words: N, sigma_x, sigma_y, length,
// x_co, y_co, addr, fallback, min_sigma_diff, enabled, then the inlined
// table [method, n, z0, dz, values[n], derivatives[n]]); interpolation
// internal/math_interpol.h:253-268, 649-665, 789-815; the last interval
// (where the reference reads one element past the table) yields 0.
ST_DEV_NOINLINE void sc_interpol( Lane& p, const LaneConst& c, cptr_t e )
{
#pragma clang fp contract( off )
    double const charge = CHARGE0_SI * c.q0;
    double const beta0 = c.beta0;
    double const beta = beta0 * p.rvv;
    double const pc0 = c.p0c * CHARGE0_SI;
    double fact_kick = c.chi * c.charge_ratio * rd( e, 3 ) * rd( e, 0 ) *
        charge * charge * ( 1.0 - beta0 * beta );
    int64_t const method = ri( e, 10 );
    int64_t const nv = ri( e, 11 );
    if( method != 255 && nv > 0 )
    {
        double const z0 = rd( e, 12 ), dz = rd( e, 13 ), z = p.zeta;
        cptr_t yv = e + 14;
        cptr_t yd = e + 14 + nv;
        double result = 0.0;
        if( ( z <= z0 + dz * nv ) && ( z >= z0 ) )
        {
            int64_t const idx = ( int64_t )( ( z - z0 ) / dz );
            if( idx + 1 < nv )
            {
                if( method == 0 )
                    result = rd( yv, idx ) + ( z - ( z0 + dz * idx ) ) * rd( yd, idx );
                else if( method == 1 )
                {
                    double const t2 = ( z0 + dz * ( idx + 1 ) - z ) / dz;
                    double const t1 = ( z - ( z0 + dz * idx ) ) / dz;
                    result = rd( yv, idx + 1 ) * t1 + rd( yv, idx ) * t2;
                    result -= rd( yd, idx ) * dz * dz * ( t2 - t2 * t2 * t2 ) / 6.0;
                    result -= rd( yd, idx + 1 ) * dz * dz * ( t1 - t1 * t1 * t1 ) / 6.0;
                }
            }
        }
        fact_kick *= result;
    }
    else fact_kick *= rd( e, 7 );
    fact_kick /= ( pc0 * beta );
    double Ex, Ey, Gx, Gy;
    gauss_field< true >( p.x - rd( e, 4 ), p.y - rd( e, 5 ), rd( e, 1 ), rd( e, 2 ),
                         rd( e, 8 ), Ex, Ey, Gx, Gy );
    p.px += fact_kick * Ex;
    p.py += fact_kick * Ey;
}
}  // namespace st

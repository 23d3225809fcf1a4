// Reverse-mode (adjoint) element maps for the differentiable tracker
// (device code).  No counterpart in the reference, which has no derivative
// of any map; each function here is the transpose of the Jacobian of the
// forward map cited next to it (track_device.h), applied to the cotangent
// of the map's output.
//
// Phase space: (x, px, y, py, zeta, delta).  rpp and rvv are functions of
// delta, rpp = 1 / (1 + delta) and rvv = (1 + delta) / D with
// D = sqrt(beta0^2 (delta^2 + 2 delta) + 1) (particles.h:5459-5497), so
// their derivatives enter every delta cotangent:
//   d rpp / d delta = -rpp^2,   d rvv / d delta = (1 - beta0^2) / D^3.
#pragma once

#include "track_device.h"

namespace st
{

// cotangent (adjoint) of the six phase-space coordinates
struct Cot
{
    double x, px, y, py, zeta, delta;
};

ST_DEV double drvv_ddelta( double const delta, double const beta0 )
{
    double const b2 = beta0 * beta0;
    double const D = sqrt( b2 * ( delta * delta + 2.0 * delta ) + 1.0 );
    return ( 1.0 - b2 ) / ( D * D * D );
}

// Drift (be_drift/track.h:45-67): x' = x + L px rpp, y' = y + L py rpp,
// zeta' = zeta + L (rvv - 1 - (xp^2 + yp^2) / 2).  p is the state before.
ST_DEV void drift_adj( const Lane& p, double const L, double const beta0, Cot& a )
{
    double const rpp = p.rpp;
    double const drpp = -rpp * rpp;
    double const xp = p.px * rpp, yp = p.py * rpp;
    double const drvv = drvv_ddelta( p.delta, beta0 );
    double const zb = a.zeta;
    a.delta += L * ( a.x * p.px * drpp + a.y * p.py * drpp +
                     zb * ( drvv - ( xp * p.px + yp * p.py ) * drpp ) );
    a.px += L * rpp * ( a.x - zb * xp );
    a.py += L * rpp * ( a.y - zb * yp );
}

// Multipole without curvature (be_multipole/track.h:38-112):
// px' = px - chi Re P(z), py' = py + chi Im P(z), z = x + i y,
// P(z) = sum_n (bal[2n] + i bal[2n+1]) z^n.  With P' the derivative,
// d(Re P)/dx = Re P', d(Im P)/dx = Im P', d(Re P)/dy = -Im P',
// d(Im P)/dy = Re P'.  p is the state before; px / py cotangents pass
// through unchanged.
ST_DEV void multipole_adj( const Lane& p, double const chi, cptr_t bal, int const order, Cot& a )
{
    double const x = p.x, y = p.y;
    // P'(z) by Horner on n b_n
    double dre = 0.0, dim = 0.0;
    for( int n = order; n >= 1; --n )
    {
        double const nre = dre * x - dim * y + n * rd( bal, 2 * n );
        double const nim = dre * y + dim * x + n * rd( bal, 2 * n + 1 );
        dre = nre;
        dim = nim;
    }
    a.x += chi * ( -a.px * dre + a.py * dim );
    a.y += chi * ( a.px * dim + a.py * dre );
}

// Cavity (be_cavity/track.h:39-67 with the energy update of
// particles.h:5255-5297): (zeta, delta) -> (zeta', delta'), zeta' =
// zeta rvv'/rvv.  The 2x2 Jacobian is formed by propagating the two unit
// tangents through the same sequence of operations, then transposed onto
// (zeta', delta') cotangents.  p is the state before; x, px, y, py are not
// touched.
ST_DEV void cavity_adj( const Lane& p, const LaneConst& c, cptr_t e, Cot& a )
{
    double const voltage = rd( e, 0 ), lag_rad = rd( e, 1 ), kfreq = rd( e, 2 );
    double const beta0 = c.beta0, b2 = beta0 * beta0;
    double const zeta = p.zeta, delta = p.delta, rvv = p.rvv;
    double const drvv = drvv_ddelta( delta, beta0 );
    // tau = zeta / (beta0 rvv)
    double const tau = zeta / ( beta0 * rvv );
    double const dtau_z = 1.0 / ( beta0 * rvv );
    double const dtau_d = -zeta * drvv / ( beta0 * rvv * rvv );
    double const phase = lag_rad - kfreq * tau;
    double const q = c.q0 * c.charge_ratio;
    double const cph = st_cos( phase );
    // dE = q V cos(phase) dphase, dphase = -k dtau
    double const dE_z = q * voltage * cph * ( -kfreq * dtau_z );
    double const dE_d = q * voltage * cph * ( -kfreq * dtau_d );
    double const energy = q * voltage * st_sin( phase );
    double const energy0 = sqrt( c.p0c * c.p0c + c.mass0 * c.mass0 );
    double const D = sqrt( b2 * ( delta * delta + 2.0 * delta ) + 1.0 );
    double const dD_d = b2 * ( delta + 1.0 ) / D;
    double const ptau_beta0 = energy / energy0 + D - 1.0;
    double const dpb_z = dE_z / energy0;
    double const dpb_d = dE_d / energy0 + dD_d;
    double const ptau = ptau_beta0 / beta0;
    double const psigma = ptau / beta0;
    double const opd_new = sqrt( ptau * ptau + 2.0 * psigma + 1.0 );     // 1 + delta'
    // d(1 + delta') = (ptau dptau + dpsigma) / (1 + delta'), dptau = dpb / beta0,
    // dpsigma = dptau / beta0
    double const ddn_z = ( ptau * dpb_z / beta0 + dpb_z / b2 ) / opd_new;
    double const ddn_d = ( ptau * dpb_d / beta0 + dpb_d / b2 ) / opd_new;
    double const den = 1.0 + ptau_beta0;
    double const rvv_new = opd_new / den;
    double const drn_z = ( ddn_z * den - opd_new * dpb_z ) / ( den * den );
    double const drn_d = ( ddn_d * den - opd_new * dpb_d ) / ( den * den );
    // zeta' = zeta rvv_new / rvv
    double const ratio = rvv_new / rvv;
    double const dzn_z = ratio + zeta * drn_z / rvv;
    double const dzn_d = zeta * ( drn_d / rvv - rvv_new * drvv / ( rvv * rvv ) );
    double const zb = a.zeta, db = a.delta;
    a.zeta = zb * dzn_z + db * ddn_z;
    a.delta = zb * dzn_d + db * ddn_d;
}

// XYShift (be_xyshift/track.h:38-55): a translation, identity Jacobian.

// SRotation (be_srotation/track.h:38-69): (x, y) and (px, py) rotated by
// the same orthogonal matrix; the adjoint applies its transpose.
ST_DEV void srotation_adj( cptr_t e, Cot& a )
{
    double const c = rd( e, 0 ), s = rd( e, 1 );
    double const xb = a.x, yb = a.y, pxb = a.px, pyb = a.py;
    a.x = c * xb - s * yb;
    a.y = s * xb + c * yb;
    a.px = c * pxb - s * pyb;
    a.py = s * pxb + c * pyb;
}

}  // namespace st

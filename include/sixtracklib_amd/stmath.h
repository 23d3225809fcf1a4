/* sixtracklib_amd/stmath.h
 *
 * Project-owned FP64 elementary functions used by every tracking map that
 * needs a transcendental (Cavity: sin; beam-beam / space-charge fields:
 * exp, sin, cos, integer and real powers).
 *
 * Why these exist: the reference tracker calls the platform libm
 * (sixtracklib/common/be_cavity/track.h:55 `sin( phase )`,
 * be_beamfields/gauss_fields.h `exp`, be_beamfields/faddeeva_cern.h
 * `exp/cos/sin/pow`).  glibc and the GPU device library round those calls
 * differently in the last bit, so no GPU tracker can be bit-identical to a
 * glibc-linked CPU tracker on lattices with cavities.  Every function here
 * is built only from IEEE-754 operations that are correctly rounded on both
 * x86-64 and gfx950 (+ - * / sqrt, rint, ldexp, frexp).  Compiled without
 * floating-point contraction, host and device therefore produce identical
 * bits, which is what lets the GPU kernels be checked bit-for-bit against
 * the CPU oracle.  Accuracy versus a correctly rounded result is ~1-2 ulp
 * (tested against libm in tests/test_stmath.py).
 *
 * Methods (all from the textbook mathematics, coefficients are exact
 * rationals 1/n! and 2/(2k+1) rounded to double):
 *   sin, cos : Cody-Waite reduction by pi/2 with a three-part constant
 *              (33 + 33 + 53 bits of pi/2) -> |r| <= pi/4, then Taylor
 *              polynomials to r^19 (sin) / r^20 (cos); truncation < 1e-19.
 *              k * (first two parts) is exact for |x| < 2^20 * pi/2; up
 *              to |x| < 2^30 the reduced argument is still good to
 *              |x| * 2^-53 absolute.  Beyond 2^30 (or for Inf / NaN) the
 *              result is NaN: a documented domain limit (libm would return
 *              a finite value; a tracking phase of 1e9 rad is not physical).
 *   exp      : k = rint(x / ln 2), two-part ln 2, Taylor to r^14 on
 *              |r| <= ln2/2, scaled with ldexp (exact, handles subnormals).
 *   log      : frexp into m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *              s = (m-1)/(m+1), series to s^23.
 *   pow      : exp(y * log x) for x > 0; integer powers by binary
 *              exponentiation (st_pow_int).
 */
#ifndef SIXTRACKLIB_AMD_STMATH_H__
#define SIXTRACKLIB_AMD_STMATH_H__

#include <math.h>

/* Every function below switches contraction off (clang only; the C build
 * of the oracle compiles with -ffp-contract=off), so that a device build
 * made with contraction still produces the same bits as the host. */
#if defined( __HIPCC__ )
    #define ST_MATH_FN __host__ __device__ static inline
#else
    #define ST_MATH_FN static inline
#endif

/* pi/2 split as 33 + 33 + 53 significant bits (first two have trailing
 * zeros so k * part is exact for |k| < 2^20). */
#define ST_PIO2_1   ( 0x1.921fb54400000p+0 )
#define ST_PIO2_2   ( 0x1.0b4611a600000p-34 )
#define ST_PIO2_3   ( 0x1.3198a2e037073p-69 )
#define ST_TWO_OVER_PI ( 0x1.45f306dc9c883p-1 )

#define ST_LN2_HI   ( 0x1.62e42fee00000p-1 )
#define ST_LN2_LO   ( 0x1.a39ef35793c76p-33 )
#define ST_INV_LN2  ( 0x1.71547652b82fep+0 )
#define ST_SQRT_HALF ( 0x1.6a09e667f3bcdp-1 )

/* sin(r) on |r| <= pi/4 : r + r^3 * (S1 + r^2 * (S2 + ...)) */
ST_MATH_FN double st_sin_kernel( double const r )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    double const z = r * r;
    double p = -0x1.2f49b46814157p-57;      /* -1/19! */
    p = p * z + 0x1.952c77030ad4ap-49;      /* +1/17! */
    p = p * z - 0x1.ae7f3e733b81fp-41;      /* -1/15! */
    p = p * z + 0x1.6124613a86d09p-33;      /* +1/13! */
    p = p * z - 0x1.ae64567f544e4p-26;      /* -1/11! */
    p = p * z + 0x1.71de3a556c734p-19;      /* +1/9!  */
    p = p * z - 0x1.a01a01a01a01ap-13;      /* -1/7!  */
    p = p * z + 0x1.1111111111111p-7;       /* +1/5!  */
    p = p * z - 0x1.5555555555555p-3;       /* -1/3!  */
    return r + r * ( z * p );
}

/* cos(r) on |r| <= pi/4 : 1 - r^2/2 + r^4 * (C2 + r^2 * (C3 + ...)),
 * with the 1 - r^2/2 step compensated for its rounding error. */
ST_MATH_FN double st_cos_kernel( double const r )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    double const z = r * r;
    double q = 0x1.e542ba4020225p-62;       /* +1/20! */
    q = q * z - 0x1.6827863b97d97p-53;      /* -1/18! */
    q = q * z + 0x1.ae7f3e733b81fp-45;      /* +1/16! */
    q = q * z - 0x1.93974a8c07c9dp-37;      /* -1/14! */
    q = q * z + 0x1.1eed8eff8d898p-29;      /* +1/12! */
    q = q * z - 0x1.27e4fb7789f5cp-22;      /* -1/10! */
    q = q * z + 0x1.a01a01a01a01ap-16;      /* +1/8!  */
    q = q * z - 0x1.6c16c16c16c17p-10;      /* -1/6!  */
    q = q * z + 0x1.5555555555555p-5;       /* +1/4!  */
    double const hz = 0.5 * z;
    double const w  = 1.0 - hz;
    return w + ( ( ( 1.0 - w ) - hz ) + ( z * z ) * q );
}

/* Reduce x by multiples of pi/2; returns r and the quadrant (k mod 4).
 * k mod 4 is formed in floating point (k/4, floor and the products are
 * exact for every finite k), so no out-of-range integer conversion can
 * occur. */
ST_MATH_FN double st_reduce_pio2( double const x, int* quadrant )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    double const k = rint( x * ST_TWO_OVER_PI );
    double r = x - k * ST_PIO2_1;
    r = r - k * ST_PIO2_2;
    r = r - k * ST_PIO2_3;
    *quadrant = ( int )( k - 4.0 * floor( 0.25 * k ) );
    return r;
}

/* Both polynomials are evaluated and the quadrant selects between them:
 * no data-dependent branch (a wave evaluates each polynomial once), and
 * the selected value is the same IEEE result a branch would produce.
 * Outside the domain (|x| >= 2^30, Inf, NaN) the result is NaN. */
ST_MATH_FN double st_sin( double const x )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    int const finite = fabs( x ) < 0x1p+30;   /* domain; false for Inf / NaN */
    int q = 0;
    double const r = st_reduce_pio2( finite ? x : 0.0, &q );
    double const s = st_sin_kernel( r ), c = st_cos_kernel( r );
    double const v = ( q & 1 ) ? c : s;
    double const res = ( q & 2 ) ? -v : v;
    return finite ? ( x == 0.0 ? x : res ) : ( double )NAN;   /* sin(-0) = -0 */
}

ST_MATH_FN double st_cos( double const x )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    int const finite = fabs( x ) < 0x1p+30;
    int q = 0;
    double const r = st_reduce_pio2( finite ? x : 0.0, &q );
    double const s = st_sin_kernel( r ), c = st_cos_kernel( r );
    double const v = ( q & 1 ) ? s : c;
    double const res = ( ( q + 1 ) & 2 ) ? -v : v;
    return finite ? res : ( double )NAN;
}

ST_MATH_FN double st_exp( double const x )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    if( x != x ) return x;
    if( x >  709.782712893384 ) return ( double )INFINITY;
    if( x < -745.1332191019412 ) return 0.0;
    double const k = rint( x * ST_INV_LN2 );
    double const r = ( x - k * ST_LN2_HI ) - k * ST_LN2_LO;
    double p = 0x1.93974a8c07c9dp-37;       /* 1/14! */
    p = p * r + 0x1.6124613a86d09p-33;      /* 1/13! */
    p = p * r + 0x1.1eed8eff8d898p-29;      /* 1/12! */
    p = p * r + 0x1.ae64567f544e4p-26;      /* 1/11! */
    p = p * r + 0x1.27e4fb7789f5cp-22;      /* 1/10! */
    p = p * r + 0x1.71de3a556c734p-19;      /* 1/9!  */
    p = p * r + 0x1.a01a01a01a01ap-16;      /* 1/8!  */
    p = p * r + 0x1.a01a01a01a01ap-13;      /* 1/7!  */
    p = p * r + 0x1.6c16c16c16c17p-10;      /* 1/6!  */
    p = p * r + 0x1.1111111111111p-7;       /* 1/5!  */
    p = p * r + 0x1.5555555555555p-5;       /* 1/4!  */
    p = p * r + 0x1.5555555555555p-3;       /* 1/3!  */
    p = p * r + 0.5;                        /* 1/2!  */
    p = r + r * ( r * p );                  /* exp(r) - 1 */
    return ldexp( 1.0 + p, ( int )k );
}

ST_MATH_FN double st_log( double const x )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    if( x != x ) return x;
    if( x < 0.0 ) return ( x - x ) / 0.0;                  /* NaN */
    if( x == 0.0 ) return -( double )INFINITY;
    if( x == ( double )INFINITY ) return x;
    int e = 0;
    double m = frexp( x, &e );                             /* [0.5, 1) */
    if( m < ST_SQRT_HALF ) { m = m + m; e = e - 1; }
    double const f = m - 1.0;
    double const s = f / ( 2.0 + f );
    double const z = s * s;
    double p = 0x1.642c8590b2164p-4;                       /* 2/23 */
    p = p * z + 0x1.8618618618618p-4;                      /* 2/21 */
    p = p * z + 0x1.af286bca1af28p-4;                      /* 2/19 */
    p = p * z + 0x1.e1e1e1e1e1e1ep-4;                      /* 2/17 */
    p = p * z + 0x1.1111111111111p-3;                      /* 2/15 */
    p = p * z + 0x1.3b13b13b13b14p-3;                      /* 2/13 */
    p = p * z + 0x1.745d1745d1746p-3;                      /* 2/11 */
    p = p * z + 0x1.c71c71c71c71cp-3;                      /* 2/9  */
    p = p * z + 0x1.2492492492492p-2;                      /* 2/7  */
    p = p * z + 0x1.999999999999ap-2;                      /* 2/5  */
    p = p * z + 0x1.5555555555555p-1;                      /* 2/3  */
    /* log m = 2 s + s z p  ;  2 s = f - s f  (exact rearrangement
     * 2s = 2f/(2+f) = f - f^2/(2+f) = f - s f) keeps the leading term sharp */
    double const logm = ( f - s * f ) + s * ( z * p );
    double const de = ( double )e;
    return de * ST_LN2_HI + ( de * ST_LN2_LO + logm );
}

/* x^n for integer n by binary exponentiation. */
ST_MATH_FN double st_pow_int( double const x, int const n )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    unsigned int m = ( n < 0 ) ? ( unsigned int )( -n ) : ( unsigned int )n;
    double base = x;
    double acc = 1.0;
    while( m != 0u )
    {
        if( m & 1u ) acc = acc * base;
        m >>= 1;
        if( m != 0u ) base = base * base;
    }
    return ( n < 0 ) ? 1.0 / acc : acc;
}

/* x^y for x >= 0 (the only use on the tracking path: q-Gaussian
 * profiles); negative x gives NaN unless y is integral. */
ST_MATH_FN double st_pow( double const x, double const y )
{
#if defined( __clang__ )
#pragma clang fp contract( off )
#endif
    if( y == 0.0 ) return 1.0;
    if( ( y == rint( y ) ) && ( fabs( y ) < 1024.0 ) )
        return st_pow_int( x, ( int )y );
    if( x == 0.0 ) return ( y > 0.0 ) ? 0.0 : ( double )INFINITY;
    return st_exp( y * st_log( x ) );
}

#endif /* SIXTRACKLIB_AMD_STMATH_H__ */

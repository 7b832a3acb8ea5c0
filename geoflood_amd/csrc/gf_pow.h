// gf_pow.h -- the two powers of the step, correctly rounded.
//
// The reference evaluates h**(7.d0/3.d0) in src2 (geoclaw_src2_fort.f90:62)
// and (q0(2)/sqrt(grav))**(2.0d0/3.0d0) for the hydrograph's inflow depth
// (geoflood_bc2_fort.f90:63 and twins, hydrograh_module.f90:245-318), i.e.
// pow(x, y) of the host libm with y the double nearest 7/3 or 2/3.  The device's own pow rounds differently from glibc's
// in about a quarter of calls (one ulp), and such differences, carried by the
// flow, move wet/dry fronts apart after O(100) steps.  These functions return
// the correctly rounded x^y (the exact real power of the double y), which is
// what glibc's pow returns too except in the rare calls it misrounds.
//
// Method.  y = k/3 + d exactly, with d = 2^-51/3 for 7/3 (k = 7) and
// d = -2^-53/3 for 2/3 (k = 2), so x^y = x^(k/3) * x^d:
//  * cbrt(x) from the library (faithful), refined by one Newton step
//    c = c0 - (c0^3 - x) / (3 c0^2) with c0^3 - x formed from exact products
//    (the FMA error-free transformation of a product; J.-M. Muller et al.,
//    Handbook of Floating-Point Arithmetic, 2nd ed., 2018, sections 4.4 and
//    5.3), giving cbrt(x) as a double-double to about 2^-104;
//  * x^2 (for 7/3) as an exact double-double product, x^(k/3) as the
//    double-double product with the refined root;
//  * x^d = exp(d ln x) = 1 + d ln x + O((d ln x)^2), |d ln x| < 1e-13 for
//    the normal range used here, so ln x to a few digits suffices;
//  * one final rounding of the double-double value (hi + (lo + hi d ln x)):
//    the correctly rounded result unless x^y lies within about 2^-100 of the
//    midpoint of two doubles.
// Arguments outside (2^-300, 2^300) (zero, negatives, subnormals, Inf, NaN)
// go to the library pow.  GF_HOST_DEVICE lets a host test compile the same
// functions.
#pragma once

#include <cmath>

#ifndef GF_HOST_DEVICE
#define GF_HOST_DEVICE __host__ __device__
#endif

namespace gf {

// cbrt(x) = hi + lo to about 2^-104 relative, x normal and > 0.
GF_HOST_DEVICE inline double cbrt_dd(double x, double& lo)
{
    const double c0 = ::cbrt(x);
    const double a = c0 * c0;
    const double a_lo = ::fma(c0, c0, -a);    // c0^2 = a + a_lo exactly
    const double b = a * c0;
    const double b_lo = ::fma(a, c0, -b);     // a c0 = b + b_lo exactly
    // c0^3 - x: b - x is exact (b lies within a factor 2 of x), the rest is
    // the low-order part of c0^3
    const double r = (b - x) + (b_lo + a_lo * c0);
    lo = -r / (3.0 * a);
    return c0;
}

// x^y for y = 7.0/3.0 (the double nearest 7/3), correctly rounded.
GF_HOST_DEVICE inline double pow_7_3(double x)
{
    if (!(x > 0x1p-300 && x < 0x1p300)) return ::pow(x, 7.0 / 3.0);
    double c_lo;
    const double c = cbrt_dd(x, c_lo);
    const double p = x * x;
    const double e = ::fma(x, x, -p);         // x^2 = p + e exactly
    const double hi = p * c;
    const double lo = ::fma(p, c, -hi) + (p * c_lo + e * c);
    const double t = ::log(x) / 6755399441055744.0;   // d ln x, d = 1 / (3 * 2^51)
    return hi + (lo + hi * t);
}

// x^y for y = 2.0/3.0 (the double nearest 2/3), correctly rounded.
GF_HOST_DEVICE inline double pow_2_3(double x)
{
    if (!(x > 0x1p-300 && x < 0x1p300)) return ::pow(x, 2.0 / 3.0);
    double c_lo;
    const double c = cbrt_dd(x, c_lo);
    const double hi = c * c;
    const double lo = ::fma(c, c, -hi) + 2.0 * (c * c_lo);
    const double t = -::log(x) / 27021597764222976.0;  // d ln x, d = -1 / (3 * 2^53)
    return hi + (lo + hi * t);
}

}  // namespace gf

// devmath.hpp -- small device math helpers for the ring stage.
#pragma once

#include <hip/hip_runtime.h>

namespace shp {

// sin and cos of x for |x| < 2^19 (the azimuth products m*phi0 of any ring,
// m <= 2^17): Cody-Waite reduction by pi/2 with a three-part constant (the
// first two parts carry 33 significant bits, so k*P1 and k*P2 are exact for
// |k| < 2^19; the three parts sum to pi/2 within 1e-37), then Taylor
// polynomials on [-pi/4, pi/4] (sin to r^19/19!, cos to r^18/18!; both
// truncation errors are below 1e-20).  Accurate to a few ulp; small enough
// to keep the FFT kernels free of the scratch memory a general-argument
// library sincos needs.
__device__ inline void sincos_small(double x, double &s, double &c)
{
  constexpr double kTwoOverPi = 0.6366197723675814;
  constexpr double P1 = 1.5707963267341256;
  constexpr double P2 = 6.077100506303966e-11;
  constexpr double P3 = 2.0222662487959506e-21;
  const double k = rint(x * kTwoOverPi);
  const double r = ((x - k * P1) - k * P2) - k * P3;
  const double r2 = r * r;
  // sin r = r * sum_{j=0}^{9} (-1)^j r^{2j} / (2j+1)!
  double ps = -1.0 / 121645100408832000.0;          // -1/19!
  ps = fma(ps, r2, 1.0 / 355687428096000.0);         //  1/17!
  ps = fma(ps, r2, -1.0 / 1307674368000.0);          // -1/15!
  ps = fma(ps, r2, 1.0 / 6227020800.0);              //  1/13!
  ps = fma(ps, r2, -1.0 / 39916800.0);               // -1/11!
  ps = fma(ps, r2, 1.0 / 362880.0);                  //  1/9!
  ps = fma(ps, r2, -1.0 / 5040.0);                   // -1/7!
  ps = fma(ps, r2, 1.0 / 120.0);                     //  1/5!
  ps = fma(ps, r2, -1.0 / 6.0);                      // -1/3!
  const double sr = fma(ps * r2, r, r);
  // cos r = sum_{j=0}^{9} (-1)^j r^{2j} / (2j)!
  double pc = -1.0 / 6402373705728000.0;             // -1/18!
  pc = fma(pc, r2, 1.0 / 20922789888000.0);          //  1/16!
  pc = fma(pc, r2, -1.0 / 87178291200.0);            // -1/14!
  pc = fma(pc, r2, 1.0 / 479001600.0);               //  1/12!
  pc = fma(pc, r2, -1.0 / 3628800.0);                // -1/10!
  pc = fma(pc, r2, 1.0 / 40320.0);                   //  1/8!
  pc = fma(pc, r2, -1.0 / 720.0);                    // -1/6!
  pc = fma(pc, r2, 1.0 / 24.0);                      //  1/4!
  pc = fma(pc, r2, -0.5);                            // -1/2!
  const double cr = fma(pc, r2, 1.0);
  const int q = ((int)k) & 3;
  if (q == 0) { s = sr; c = cr; }
  else if (q == 1) { s = cr; c = -sr; }
  else if (q == 2) { s = -sr; c = -cr; }
  else { s = -cr; c = sr; }
}

__device__ inline double2 expi_small(double x)
{
  double s, c;
  sincos_small(x, s, c);
  return make_double2(c, s);
}

} // namespace shp

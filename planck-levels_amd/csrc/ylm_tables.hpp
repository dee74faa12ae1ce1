// ylm_tables.hpp -- host-side recurrence tables of the Y_lm generators, one m
// at a time.  The engine's device tables (engine.cpp) and the exported
// sharp_Ylmgen_* functions (ylmgen_abi.cpp) are both built with these, so
// they hold the same values.  The formulas and their order of operations are
// libsharp's (libsharp/sharp_ylmgen_c.c:44-232): the coefficients equal the
// reference's up to the regrouping of products its -ffast-math build allows
// (2 ulp).
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "shp_internal.hpp"

namespace shp {

// v * 2^(800 scale) kept with |v| in [xfmax 2^-800, xfmax]
// (libsharp/sharp_ylmgen_c.c:37-42)
inline void ylm_normalize(double &val, int &scale, double xfmax)
{
  while (std::fabs(val) > xfmax) { val *= kFSmall; ++scale; }
  if (val != 0.)
    while (std::fabs(val) < xfmax * kFSmall) { val *= kFBig; --scale; }
}

// sqrt(k) and 1/sqrt(k) (0 for k = 0) for k < 2 lmax + 5
// (libsharp/sharp_ylmgen_c.c:69-75)
inline void ylm_roots(int lmax, std::vector<double> &root, std::vector<double> &iroot)
{
  root.resize(2 * (size_t)lmax + 5);
  iroot.resize(root.size());
  for (size_t k = 0; k < root.size(); ++k) {
    root[k] = std::sqrt((double)k);
    iroot[k] = (k == 0) ? 0. : 1. / root[k];
  }
}

// lambda_mm prefactors mfac[0..mmax] (libsharp/sharp_ylmgen_c.c:63-67)
inline void ylm_mfac(int mmax, double *mfac)
{
  mfac[0] = 0.2820947917738781434740397257803862929220;   // 1/sqrt(4 pi)
  for (int m = 1; m <= mmax; ++m) mfac[m] = mfac[m - 1] * std::sqrt((2 * m + 1.) / (2 * m));
}

// Spin-0 coefficients of one m: lambda_{l+1} = a_l cth lambda_l - b_l
// lambda_{l-1}, {a_l, b_l} for l = m..lmax at rf[2 (l-m)], rf[2 (l-m) + 1]
// (libsharp/sharp_ylmgen_c.c:160-171)
inline void ylm_spin0_coeffs(int m, int lmax, const double *root, const double *iroot,
                             double *rf)
{
  rf[0] = root[2 * m + 3];
  rf[1] = 0.;
  for (int l = m + 1; l <= lmax; ++l) {
    const double tmp = root[2 * l + 3] * iroot[l + 1 + m] * iroot[l + 1 - m];
    rf[2 * (l - m)] = tmp * root[2 * l + 1];
    rf[2 * (l - m) + 1] = tmp * root[l + m] * root[l - m] * iroot[2 * l - 1];
  }
}

// 1/k (0 for k = 0), sqrt(1/(k+1)) and sqrt(k/(k+1)) of the spin recurrences
// (libsharp/sharp_ylmgen_c.c:82-92)
struct SpinFactors {
  std::vector<double> inv, flm1, flm2;
  explicit SpinFactors(int lmax) : inv(lmax + 1), flm1(2 * (size_t)lmax + 1), flm2(flm1.size())
  {
    inv[0] = 0;
    for (int k = 1; k <= lmax; ++k) inv[k] = 1. / k;
    for (size_t k = 0; k < flm1.size(); ++k) {
      flm1[k] = std::sqrt(1. / (k + 1.));
      flm2[k] = std::sqrt(k / (k + 1.));
    }
  }
};

// Spin-s coefficients of one m: entry l+1 = {f0, f1, f2} for l = max(m,s)
// .. lmax-1 at fx[3 (l+1-base) + 0..2], the recurrence lambda_{l+1} =
// (cth f0 -+ f1 f0) lambda_l - f2 lambda_{l-1} of each chain
// (libsharp/sharp_ylmgen_c.c:180-194)
inline void ylm_spin_coeffs(int m, int s, int lmax, const double *flm1, const double *flm2,
                            const double *inv, double *fx, int base)
{
  const int mhi = std::max(m, s);
  for (int l = mhi; l < lmax; ++l) {
    double t = flm1[l + m] * flm1[l - m] * flm1[l + s] * flm1[l - s];
    const double lt = 2 * l + 1, l1 = l + 1;
    double *e = fx + 3 * (l + 1 - base);
    e[0] = l1 * lt * t;
    e[1] = m * s * inv[l] * inv[l + 1];
    t = flm2[l + m] * flm2[l - m] * flm2[l + s] * flm2[l - s];
    e[2] = t * l1 * inv[l];
  }
}

// Spin-s start-value prefactors (2 mhi)! / ((mhi+mlo)! (mhi-mlo)!) under a
// square root, as value and scale, for m = 0..mmax
// (libsharp/sharp_ylmgen_c.c:93-117)
inline void ylm_spin_prefactors(int lmax, int mmax, int s, double *prefac, int *fscale)
{
  std::vector<double> fac(2 * (size_t)lmax + 1);
  std::vector<int> facscale(fac.size());
  fac[0] = 1;
  facscale[0] = 0;
  for (size_t k = 1; k < fac.size(); ++k) {
    fac[k] = fac[k - 1] * std::sqrt((double)k);
    facscale[k] = facscale[k - 1];
    ylm_normalize(fac[k], facscale[k], kFBigHalf);
  }
  for (int m = 0; m <= mmax; ++m) {
    int mlo = s, mhi = m;
    if (mhi < mlo) std::swap(mhi, mlo);
    double tfac = fac[2 * mhi] / fac[mhi + mlo];
    int tscale = facscale[2 * mhi] - facscale[mhi + mlo];
    ylm_normalize(tfac, tscale, kFBigHalf);
    tfac /= fac[mhi - mlo];
    tscale -= facscale[mhi - mlo];
    ylm_normalize(tfac, tscale, kFBigHalf);
    prefac[m] = tfac;
    fscale[m] = tscale;
  }
}

// norm_l for l = 0..lmax (libsharp/sharp_ylmgen_c.c:200-232): 1 for spin 0;
// for spin s, 0 below l = s and sign/2 sqrt((2l+1)/(4 pi)) above (sign -1
// for s > 0, flipped for odd s); deriv: 1/2 sqrt(l(l+1)(2l+1)/(4 pi))
inline void ylm_norm(int lmax, int spin, bool deriv, double *n)
{
  const double pi = 3.141592653589793238462643383279502884197;
  if (deriv) {
    for (int l = 0; l <= lmax; ++l)
      n[l] = (l < 1) ? 0. : 0.5 * std::sqrt(l * (l + 1.) * (2 * l + 1.) / (4 * pi));
  } else if (spin == 0) {
    for (int l = 0; l <= lmax; ++l) n[l] = 1.;
  } else {
    double spinsign = (spin > 0) ? -1.0 : 1.0;
    spinsign = (spin & 1) ? -spinsign : spinsign;
    for (int l = 0; l <= lmax; ++l)
      n[l] = (l < spin) ? 0. : spinsign * 0.5 * std::sqrt((2 * l + 1) / (4 * pi));
  }
}

} // namespace shp

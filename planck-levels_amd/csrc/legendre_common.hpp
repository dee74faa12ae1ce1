// legendre_common.hpp -- device helpers shared by the FP64 and the mixed-
// precision Legendre kernels: exponent-scaled start values and the spin-0 /
// spin-s recurrence steps, in libsharp's operation order
// (libsharp/sharp_core_inc.c:73-293).
#pragma once

#include <hip/hip_runtime.h>
#include "shp_internal.hpp"

namespace shp {

// Maps a linear workgroup id to (major, minor) so that the nminor workgroups
// of one major index land on the same XCD (the hardware deals workgroup ids
// round-robin over the 8 XCDs, each with its own L2): workgroups that read
// the same rows share one L2.  False for the padding ids of the last round.
__device__ inline bool xcd_map(int id, int nmajor, int nminor, int &major, int &minor)
{
  const int xcd = id & 7, k = id >> 3;
  major = (k / nminor) * 8 + xcd;
  minor = k % nminor;
  return major < nmajor;
}

static inline int xcd_grid(int nmajor, int nminor)
{
  return ((nmajor + 7) / 8) * 8 * nminor;
}

// Brings |val| into [maxval*2^-800, maxval].  A finite double needs at most
// three 800-binade steps; the bound also ends the loop for non-finite input.
__device__ inline void normalize(double &val, int &scale, double maxval)
{
  for (int it = 0; it < 4 && fabs(val) > maxval; ++it) { val *= kFSmall; ++scale; }
  if (val != 0.)
    for (int it = 0; it < 4 && fabs(val) < maxval * kFSmall; ++it) {
      val *= kFBig;
      --scale;
    }
}

// val^npow with a separate exponent (libsharp/sharp_core_inc.c:97-126)
__device__ inline void mypow(double val, int npow, double &resd, int &ress)
{
  int scale = 0, scaleint = 0;
  double res = 1.;
  normalize(val, scaleint, kFBigHalf);
  do {
    if (npow & 1) {
      res *= val;
      scale += scaleint;
      normalize(res, scale, kFBigHalf);
    }
    val *= val;
    scaleint += scaleint;
    normalize(val, scaleint, kFBigHalf);
  } while (npow >>= 1);
  resd = res;
  ress = scale;
}

__device__ inline double corfac(int scale)
{
  return (scale < 0) ? 0. : ((scale == 0) ? 1. : kFBig);
}

__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }
__device__ inline bool wave_all(bool p) { return __ballot(!p) == 0ull; }

struct SpinStart {
  int mhi, cosPow, sinPow, preMinus_p, preMinus_m;
};

__device__ inline SpinStart spin_start(int m, int s)
{
  SpinStart r;
  const int mhi = (m < s) ? s : m;
  r.mhi = mhi;
  r.preMinus_p = r.preMinus_m = 0;
  if (mhi == m) {
    r.cosPow = mhi + s;
    r.sinPow = mhi - s;
    r.preMinus_p = r.preMinus_m = ((mhi - s) & 1);
  } else {
    r.cosPow = mhi + m;
    r.sinPow = mhi - m;
    r.preMinus_m = ((mhi + m) & 1);
  }
  return r;
}

// start values of lambda_plus / lambda_minus at l=mhi with their scales
// (libsharp/sharp_core_inc.c:223-273)
__device__ inline void spin_init(double cth, double sth, int s, const SpinStart &st,
                                 double prefac, int fscale, double &rec2p,
                                 double &rec2m, int &scalep, int &scalem)
{
  double cth2 = fmax(sqrt((1. + cth) * 0.5), 1e-15);
  double sth2 = fmax(sqrt((1. - cth) * 0.5), 1e-15);
  if (sth < 0. && cth < 0.) cth2 = -cth2;
  if (sth < 0. && cth > 0.) sth2 = -sth2;
  double ccp, ssp, csp, scp;
  int ccps, ssps, csps, scps;
  mypow(cth2, st.cosPow, ccp, ccps);
  mypow(sth2, st.sinPow, ssp, ssps);
  mypow(cth2, st.sinPow, csp, csps);
  mypow(sth2, st.cosPow, scp, scps);
  rec2p = prefac * ccp; scalep = fscale + ccps;
  rec2m = prefac * csp; scalem = fscale + csps;
  normalize(rec2m, scalem, kFBigHalf);
  normalize(rec2p, scalep, kFBigHalf);
  rec2p *= ssp; scalep += ssps;
  rec2m *= scp; scalem += scps;
  if (st.preMinus_p) rec2p = -rec2p;
  if (st.preMinus_m) rec2m = -rec2m;
  if (s & 1) rec2p = -rec2p;
  normalize(rec2m, scalem, kFTol);
  normalize(rec2p, scalep, kFTol);
}

// spin-s recurrence coefficients of one l: libsharp's fx[l] = {f0, f1, f2}
// (libsharp/sharp_ylmgen_c.c:174-183) with f1 stored premultiplied, g = f0*f1
struct Fx { double f0, g, f2; };

// x_{l+1} = (cth -+ f1) f0 y_l - f2 x_{l-1} (libsharp/sharp_core_inc.c:
// 209-221), the factor (cth -+ f1) f0 = cth f0 -+ g as one FMA
__device__ inline void rec_step(double &rxp, double &rxm, double ryp, double rym,
                                double cth, const Fx &f)
{
  rxp = fma(fma(cth, f.f0, -f.g), ryp, -(f.f2 * rxp));
  rxm = fma(fma(cth, f.f0, f.g), rym, -(f.f2 * rxm));
}

// rec_step on the table whose f2 are all 1: the recurrence of lambda_l / s_l
// with s_l = f2_l s_{l-2} (s = 1 at the first two l), whose factor s_l the
// a_lm gather and scatter apply instead (the mixed-precision kernels)
__device__ inline void rec_step_unit(double &rxp, double &rxm, double ryp, double rym,
                                     double cth, const Fx &f)
{
  rxp = fma(fma(cth, f.f0, -f.g), ryp, -rxp);
  rxm = fma(fma(cth, f.f0, f.g), rym, -rxm);
}

__device__ inline bool rescale1(double &l1, double &l2, int &scale)
{
  if (fabs(l2) > kFTol) {
    l1 *= kFSmall; l2 *= kFSmall; ++scale;
    return true;
  }
  return false;
}

} // namespace shp

// sharp_abi.cpp -- host-side half of the drop-in C interface: a_lm and
// geometry info objects, the standard geometries, the per-ring m cutoff and
// the error convention.  The GPU engine (engine.hip) consumes the objects
// built here; callers may also build or edit them by hand, as libsharp's own
// tests do.
//
// Reference behaviour followed:
//   libsharp/sharp.c:66-78        sharp_get_mlim
//   libsharp/sharp.c:122-233      info objects, ring pairing and sort order
//   libsharp/sharp_geomhelpers.c  HEALPix / Gauss / Fejer / CC / MW grids
//   libsharp/sharp_almhelpers.c   triangular / rectangular a_lm layouts
//   c_utils/c_utils.c:35-39       "file, line (func):\nmsg" then exit(1)

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "sharp_lowlevel.h"
#include "sharp_geomhelpers.h"
#include "sharp_almhelpers.h"
#include "shp_internal.hpp"

namespace shp {

[[noreturn]] void fail(const char *file, int line, const char *func,
                       const char *msg)
{
  std::fprintf(stderr, "%s, %i (%s):\n%s\n", file, line, func, msg);
  std::fflush(stderr);
  std::exit(1);
}

// Allocation for info objects: callers free them only through the matching
// sharp_destroy_* function, so plain malloc/free is sufficient.
template <typename T> static T *xalloc(size_t n)
{
  if (n == 0) return nullptr;
  T *p = static_cast<T *>(std::malloc(n * sizeof(T)));
  if (!p) SHP_FAIL("malloc() failed");
  return p;
}

static const double kPi = 3.141592653589793238462643383279502884197;

// Unnormalised backward real DFT in fftpack storage (r0, r1, i1, r2, i2, ...)
// used only for the Fejer / Clenshaw-Curtis quadrature weights.  A direct
// O(n^2) sum with an exact twiddle table; n is at most a few thousand.
static void real_dft_backward(int n, double *data)
{
  std::vector<double> c(n), s(n), out(n);
  for (int k = 0; k < n; ++k) {
    c[k] = std::cos(2 * kPi * k / n);
    s[k] = std::sin(2 * kPi * k / n);
  }
  for (int j = 0; j < n; ++j) {
    double acc = data[0];
    for (int k = 1; 2 * k < n; ++k) {
      int idx = int((long long)k * j % n);
      acc += 2. * (data[2 * k - 1] * c[idx] - data[2 * k] * s[idx]);
    }
    if ((n & 1) == 0) acc += data[n - 1] * ((j & 1) ? -1. : 1.);
    out[j] = acc;
  }
  std::memcpy(data, out.data(), sizeof(double) * n);
}

} // namespace shp

using namespace shp;

extern "C" {

// ---------------------------------------------------------------------------
// a_lm info
// ---------------------------------------------------------------------------
void sharp_make_general_alm_info(int lmax, int nm, int stride, const int *mval,
                                 const ptrdiff_t *mvstart, int flags,
                                 sharp_alm_info **alm_info)
{
  sharp_alm_info *info = xalloc<sharp_alm_info>(1);
  info->lmax = lmax;
  info->nm = nm;
  info->mval = xalloc<int>(nm);
  info->mvstart = xalloc<ptrdiff_t>(nm);
  info->stride = stride;
  info->flags = flags;
  for (int i = 0; i < nm; ++i) {
    info->mval[i] = mval[i];
    info->mvstart[i] = mvstart[i];
  }
  *alm_info = info;
}

void sharp_make_alm_info(int lmax, int mmax, int stride, const ptrdiff_t *mstart,
                         sharp_alm_info **alm_info)
{
  std::vector<int> mval(mmax + 1);
  for (int m = 0; m <= mmax; ++m) mval[m] = m;
  sharp_make_general_alm_info(lmax, mmax + 1, stride, mval.data(), mstart, 0,
                              alm_info);
}

ptrdiff_t sharp_alm_index(const sharp_alm_info *self, int l, int mi)
{
  if (self->flags & SHARP_PACKED)
    SHP_FAIL("sharp_alm_index not applicable with SHARP_PACKED alms");
  return self->mvstart[mi] + self->stride * l;
}

void sharp_destroy_alm_info(sharp_alm_info *info)
{
  if (!info) return;
  std::free(info->mval);
  std::free(info->mvstart);
  std::free(info);
}

void sharp_make_triangular_alm_info(int lmax, int mmax, int stride,
                                    sharp_alm_info **alm_info)
{
  sharp_alm_info *info = xalloc<sharp_alm_info>(1);
  info->lmax = lmax;
  info->nm = mmax + 1;
  info->mval = xalloc<int>(mmax + 1);
  info->mvstart = xalloc<ptrdiff_t>(mmax + 1);
  info->stride = stride;
  info->flags = 0;
  const ptrdiff_t tval = 2 * (ptrdiff_t)lmax + 1;
  for (ptrdiff_t m = 0; m <= mmax; ++m) {
    info->mval[m] = int(m);
    info->mvstart[m] = stride * ((m * (tval - m)) >> 1);
  }
  *alm_info = info;
}

void sharp_make_rectangular_alm_info(int lmax, int mmax, int stride,
                                     sharp_alm_info **alm_info)
{
  sharp_alm_info *info = xalloc<sharp_alm_info>(1);
  info->lmax = lmax;
  info->nm = mmax + 1;
  info->mval = xalloc<int>(mmax + 1);
  info->mvstart = xalloc<ptrdiff_t>(mmax + 1);
  info->stride = stride;
  info->flags = 0;
  for (ptrdiff_t m = 0; m <= mmax; ++m) {
    info->mval[m] = int(m);
    info->mvstart[m] = stride * m * (lmax + 1);
  }
  *alm_info = info;
}

// ---------------------------------------------------------------------------
// geometry info
// ---------------------------------------------------------------------------
void sharp_make_geom_info(int nrings, const int *nph, const ptrdiff_t *ofs,
                          const int *stride, const double *phi0,
                          const double *theta, const double *wgt,
                          sharp_geom_info **geom_info)
{
  sharp_geom_info *info = xalloc<sharp_geom_info>(1);
  std::vector<sharp_ringinfo> ri(nrings);
  info->pair = xalloc<sharp_ringpair>(nrings);
  info->npairs = 0;
  info->nphmax = 0;
  for (int i = 0; i < nrings; ++i) {
    ri[i].theta = theta[i];
    ri[i].cth = std::cos(theta[i]);
    ri[i].sth = std::sin(theta[i]);
    ri[i].weight = wgt ? wgt[i] : 1.;
    ri[i].phi0 = phi0[i];
    ri[i].ofs = ofs[i];
    ri[i].stride = stride[i];
    ri[i].nph = nph[i];
    info->nphmax = std::max(info->nphmax, nph[i]);
  }
  // Rings ordered by sin(theta); neighbours with cos(theta) of opposite sign
  // (to 1e-12 relative) form a pair with the northern ring first.
  // (libsharp/sharp.c:204-223; qsort there is not stable either, ties keep
  // the order of a stable sort here, which equals qsort's for distinct keys.)
  std::stable_sort(ri.begin(), ri.end(),
                   [](const sharp_ringinfo &a, const sharp_ringinfo &b) {
                     return a.sth < b.sth;
                   });
  for (int pos = 0; pos < nrings; ++pos) {
    sharp_ringpair &p = info->pair[info->npairs++];
    p.r1 = ri[pos];
    if (pos < nrings - 1 &&
        std::fabs(ri[pos].cth + ri[pos + 1].cth) < 1e-12 * std::fabs(ri[pos + 1].cth)) {
      if (ri[pos].cth > 0)
        p.r2 = ri[pos + 1];
      else {
        p.r1 = ri[pos + 1];
        p.r2 = ri[pos];
      }
      ++pos;
    } else {
      std::memset(&p.r2, 0, sizeof(p.r2));
      p.r2.nph = -1;
    }
  }
  // Pair order: by r1.nph, then r1.phi0, then descending r1.cth
  // (libsharp/sharp.c:127-135,226).
  std::stable_sort(info->pair, info->pair + info->npairs,
                   [](const sharp_ringpair &a, const sharp_ringpair &b) {
                     if (a.r1.nph != b.r1.nph) return a.r1.nph < b.r1.nph;
                     if (a.r1.phi0 != b.r1.phi0) return a.r1.phi0 < b.r1.phi0;
                     return a.r1.cth > b.r1.cth;
                   });
  *geom_info = info;
}

void sharp_destroy_geom_info(sharp_geom_info *info)
{
  if (!info) return;
  std::free(info->pair);
  std::free(info);
}

void sharp_make_weighted_healpix_geom_info(int nside, int stride,
                                           const double *weight,
                                           sharp_geom_info **geom_info)
{
  const ptrdiff_t npix = (ptrdiff_t)nside * nside * 12;
  const ptrdiff_t ncap = 2 * (ptrdiff_t)nside * (nside - 1);
  const int nrings = 4 * nside - 1;
  std::vector<double> theta(nrings), wgt(nrings), phi0(nrings);
  std::vector<int> nph(nrings), str(nrings);
  std::vector<ptrdiff_t> ofs(nrings);
  for (int m = 0; m < nrings; ++m) {
    const int ring = m + 1;
    const ptrdiff_t nr = (ring > 2 * nside) ? 4 * nside - ring : ring;
    str[m] = stride;
    if (nr < nside) { // polar cap: 4*nr pixels
      theta[m] = 2 * std::asin(nr / (std::sqrt(6.) * nside));
      nph[m] = 4 * int(nr);
      phi0[m] = kPi / nph[m];
      ofs[m] = 2 * nr * (nr - 1) * stride;
    } else { // equatorial belt: 4*nside pixels, every other ring shifted
      const double fact1 = (8. * nside) / npix;
      theta[m] = std::acos((2 * nside - nr) * fact1);
      nph[m] = 4 * nside;
      phi0[m] = ((nr - nside) & 1) ? 0 : kPi / nph[m];
      ofs[m] = (ncap + (nr - nside) * nph[m]) * stride;
    }
    if (nr != ring) { // southern hemisphere
      theta[m] = kPi - theta[m];
      ofs[m] = (npix - nph[m]) * stride - ofs[m];
    }
    wgt[m] = 4. * kPi / npix * (weight ? weight[nr - 1] : 1.);
  }
  sharp_make_geom_info(nrings, nph.data(), ofs.data(), str.data(), phi0.data(),
                       theta.data(), wgt.data(), geom_info);
}

static double one_minus_x2(double x)
{
  return (std::fabs(x) > 0.1) ? (1. + x) * (1. - x) : 1. - x * x;
}

// Gauss-Legendre nodes and weights by Newton iteration on P_n, with the
// iteration and stopping rule of libsharp/sharp_geomhelpers.c:105-156.
static void gauss_legendre(int n, double *x, double *w)
{
  const double eps = 3e-14;
  const int m = (n + 1) >> 1;
  const double t0 = 1 - (1 - 1. / n) / (8. * n * n);
  const double t1 = 1. / (4. * n + 2.);
  for (int i = 1; i <= m; ++i) {
    double x0 = std::cos(kPi * ((i << 2) - 1) * t1) * t0;
    bool dobreak = false;
    int j = 0;
    double dpdx;
    for (;;) {
      double P_1 = 1.0, P0 = x0;
      for (int k = 2; k <= n; k++) {
        const double P_2 = P_1;
        P_1 = P0;
        P0 = x0 * P_1 + (k - 1.) / k * (x0 * P_1 - P_2);
      }
      dpdx = (P_1 - x0 * P0) * n / one_minus_x2(x0);
      const double x1 = x0 - P0 / dpdx;
      const double dx = x0 - x1;
      x0 = x1;
      if (dobreak) break;
      if (std::fabs(dx) <= eps) dobreak = true;
      if (++j >= 100) SHP_FAIL("convergence problem");
    }
    x[i - 1] = -x0;
    x[n - i] = x0;
    w[i - 1] = w[n - i] = 2. / (one_minus_x2(x0) * dpdx * dpdx);
  }
}

static void make_iso(int nrings, int nphi, double phi0, int stride_lon,
                     int stride_lat, const double *theta, const double *wgt,
                     sharp_geom_info **geom_info)
{
  std::vector<int> nph(nrings, nphi), str(nrings, stride_lon);
  std::vector<double> ph(nrings, phi0);
  std::vector<ptrdiff_t> ofs(nrings);
  for (int m = 0; m < nrings; ++m) ofs[m] = (ptrdiff_t)m * stride_lat;
  sharp_make_geom_info(nrings, nph.data(), ofs.data(), str.data(), ph.data(),
                       theta, wgt, geom_info);
}

void sharp_make_gauss_geom_info(int nrings, int nphi, double phi0,
                                int stride_lon, int stride_lat,
                                sharp_geom_info **geom_info)
{
  std::vector<double> theta(nrings), w(nrings);
  gauss_legendre(nrings, theta.data(), w.data());
  for (int m = 0; m < nrings; ++m) {
    theta[m] = std::acos(-theta[m]);
    w[m] *= 2 * kPi / nphi;
  }
  make_iso(nrings, nphi, phi0, stride_lon, stride_lat, theta.data(), w.data(),
           geom_info);
}

// Fejer / Clenshaw-Curtis weights after Waldvogel (2006), BIT 46:195,
// as used by libsharp/sharp_geomhelpers.c:192-334.
void sharp_make_fejer1_geom_info(int nrings, int ppring, double phi0,
                                 int stride_lon, int stride_lat,
                                 sharp_geom_info **geom_info)
{
  std::vector<double> theta(nrings), w(nrings);
  w[0] = 2.;
  for (int k = 1; k <= (nrings - 1) / 2; ++k) {
    w[2 * k - 1] = 2. / (1. - 4. * k * k) * std::cos((k * kPi) / nrings);
    w[2 * k] = 2. / (1. - 4. * k * k) * std::sin((k * kPi) / nrings);
  }
  if ((nrings & 1) == 0) w[nrings - 1] = 0.;
  real_dft_backward(nrings, w.data());
  for (int m = 0; m < (nrings + 1) / 2; ++m) {
    theta[m] = kPi * (m + 0.5) / nrings;
    theta[nrings - 1 - m] = kPi - theta[m];
    w[m] = w[nrings - 1 - m] = w[m] * 2 * kPi / (nrings * (double)ppring);
  }
  make_iso(nrings, ppring, phi0, stride_lon, stride_lat, theta.data(), w.data(),
           geom_info);
}

void sharp_make_cc_geom_info(int nrings, int ppring, double phi0,
                             int stride_lon, int stride_lat,
                             sharp_geom_info **geom_info)
{
  const int n = nrings - 1;
  std::vector<double> theta(nrings), w(nrings, 0.);
  const double dw = -1. / (n * n - 1. + (n & 1));
  w[0] = 2. + dw;
  for (int k = 1; k <= (n / 2 - 1); ++k) w[2 * k - 1] = 2. / (1. - 4. * k * k) + dw;
  w[2 * (n / 2) - 1] = (n - 3.) / (2 * (n / 2) - 1) - 1. - dw * ((2 - (n & 1)) * n - 1);
  real_dft_backward(n, w.data());
  w[n] = w[0];
  for (int m = 0; m < (nrings + 1) / 2; ++m) {
    theta[m] = kPi * m / (nrings - 1.);
    if (theta[m] < 1e-15) theta[m] = 1e-15;
    theta[nrings - 1 - m] = kPi - theta[m];
    w[m] = w[nrings - 1 - m] = w[m] * 2 * kPi / (n * (double)ppring);
  }
  make_iso(nrings, ppring, phi0, stride_lon, stride_lat, theta.data(), w.data(),
           geom_info);
}

void sharp_make_fejer2_geom_info(int nrings, int ppring, double phi0,
                                 int stride_lon, int stride_lat,
                                 sharp_geom_info **geom_info)
{
  const int n = nrings + 1;
  std::vector<double> theta(nrings), w(n, 0.);
  w[0] = 2.;
  for (int k = 1; k <= (n / 2 - 1); ++k) w[2 * k - 1] = 2. / (1. - 4. * k * k);
  w[2 * (n / 2) - 1] = (n - 3.) / (2 * (n / 2) - 1) - 1.;
  real_dft_backward(n, w.data());
  for (int m = 0; m < nrings; ++m) w[m] = w[m + 1];
  for (int m = 0; m < (nrings + 1) / 2; ++m) {
    theta[m] = kPi * (m + 1) / (nrings + 1.);
    theta[nrings - 1 - m] = kPi - theta[m];
    w[m] = w[nrings - 1 - m] = w[m] * 2 * kPi / (n * (double)ppring);
  }
  make_iso(nrings, ppring, phi0, stride_lon, stride_lat, theta.data(), w.data(),
           geom_info);
}

void sharp_make_mw_geom_info(int nrings, int ppring, double phi0,
                             int stride_lon, int stride_lat,
                             sharp_geom_info **geom_info)
{
  std::vector<double> theta(nrings);
  for (int m = 0; m < nrings; ++m) {
    theta[m] = kPi * (2. * m + 1.) / (2. * nrings - 1.);
    if (theta[m] > kPi - 1e-15) theta[m] = kPi - 1e-15;
  }
  make_iso(nrings, ppring, phi0, stride_lon, stride_lat, theta.data(), nullptr,
           geom_info);
}

// ---------------------------------------------------------------------------
// per-ring m cutoff: m is skipped for a ring pair when m exceeds
// lmax*sin(theta) + max(100, 0.01*lmax), generalised to spin by the positive
// root of the quadratic (libsharp/sharp.c:66-78).
// ---------------------------------------------------------------------------
int sharp_get_mlim(int lmax, int spin, double sth, double cth)
{
  double ofs = lmax * 0.01;
  if (ofs < 100.) ofs = 100.;
  const double b = -2 * spin * std::fabs(cth);
  const double t1 = lmax * sth + ofs;
  const double c = (double)spin * spin - t1 * t1;
  const double discr = b * b - 4 * c;
  if (discr <= 0) return lmax;
  double res = (-b + std::sqrt(discr)) / 2.;
  if (res > lmax) res = lmax;
  return (int)(res + 0.5);
}

// CPU vector-blocking knobs of libsharp: accepted, without effect on the GPU.
void sharp_set_chunksize_min(int) {}
void sharp_set_nchunks_max(int) {}
int sharp_get_nv_max(void) { return 6; }
int sharp_nv_oracle(sharp_jobtype type, int spin, int ntrans)
{
  if ((int)type >= 5) SHP_FAIL("bad type");
  if (ntrans <= 0) SHP_FAIL("bad number of simultaneous transforms");
  if (spin < 0) SHP_FAIL("bad spin");
  return 1;
}

} // extern "C"

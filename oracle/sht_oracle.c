/*
 * sht_oracle.c -- plain-C restatement of libsharp's sharp_execute path.
 *
 * TEST INFRASTRUCTURE ONLY (see sht_oracle.h).  Never linked into the product.
 *
 * Structure of one call (libsharp/sharp.c:762-880):
 *   1. job normalisation           orc_execute            sharp.c:840-867
 *   2. output clearing             clear_outputs          sharp.c:355-453
 *   3. map -> phase (analysis)     ring_to_phase          sharp.c:305-353,631-652
 *   4. per (m, ring pair):
 *        lambda_lm by recurrence   orc_ylm_spin0/_spin    sharp_core_inc.c:73-293,
 *                                                         sharp_core_inc2.c:164-280,432-519
 *        accumulate                synth_pair/anal_pair   sharp_core_inc2.c:624-801
 *   5. phase -> map (synthesis)    phase_to_ring          sharp.c:253-303,653-669
 * Everything is scalar (one "lane" per ring pair); the ring FFTs of
 * libfftpack are replaced by direct DFTs, which compute the same sums.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "sht_oracle.h"

typedef struct { double r, i; } cplx;

static const double pi = 3.141592653589793238462643383279502884197;
static const double sqrt_one_half = 0.707106781186547572737310929369;
static const double sqrt_two = 1.414213562373095145474621858739;

/* exponent-scaling constants, libsharp/sharp_ylmgen_c.h:39-42 */
static const double fbig = 0x1p+800, fsmall = 0x1p-800;
static const double ftol = 0x1p-60, fbighalf = 0x1p+400;

static void fail (const char *func, const char *msg)
  {
  fprintf(stderr, "sht_oracle (%s):\n%s\n", func, msg);
  exit(1);
  }

static void *xmalloc (size_t sz)
  {
  void *res = malloc(sz ? sz : 1);
  if (!res) fail("xmalloc", "malloc() failed");
  return res;
  }

static int approx (double a, double b, double eps)
  { return fabs(a-b) < eps*fabs(b); }

/* ------------------------------------------------------------------------
   geometry, libsharp/sharp.c:122-135,179-227
   ------------------------------------------------------------------------ */
static int cmp_ring (const void *xa, const void *xb)
  {
  const orc_ringinfo *a = xa, *b = xb;
  return (a->sth < b->sth) ? -1 : (a->sth > b->sth) ? 1 : 0;
  }
static int cmp_pair (const void *xa, const void *xb)
  {
  const orc_ringpair *a = xa, *b = xb;
  if (a->r1.nph == b->r1.nph)
    return (a->r1.phi0 < b->r1.phi0) ? -1 :
      ((a->r1.phi0 > b->r1.phi0) ? 1 : (a->r1.cth > b->r1.cth ? -1 : 1));
  return (a->r1.nph < b->r1.nph) ? -1 : 1;
  }

void orc_make_geom_info (int nrings, const int *nph, const ptrdiff_t *ofs,
  const int *stride, const double *phi0, const double *theta,
  const double *wgt, orc_geom_info **geom_info)
  {
  orc_geom_info *info = xmalloc(sizeof(*info));
  orc_ringinfo *ri = xmalloc(sizeof(orc_ringinfo)*(size_t)nrings);
  info->pair = xmalloc(sizeof(orc_ringpair)*(size_t)nrings);
  info->npairs = 0;
  info->nphmax = 0;
  for (int i=0; i<nrings; ++i)
    {
    ri[i].theta = theta[i];
    ri[i].cth = cos(theta[i]);
    ri[i].sth = sin(theta[i]);
    ri[i].weight = wgt ? wgt[i] : 1.;
    ri[i].phi0 = phi0[i];
    ri[i].ofs = ofs[i];
    ri[i].stride = stride[i];
    ri[i].nph = nph[i];
    if (info->nphmax < nph[i]) info->nphmax = nph[i];
    }
  qsort(ri, (size_t)nrings, sizeof(orc_ringinfo), cmp_ring);
  for (int pos=0; pos<nrings; ++pos)
    {
    orc_ringpair *p = &info->pair[info->npairs++];
    p->r1 = ri[pos];
    if ((pos<nrings-1) && approx(ri[pos].cth, -ri[pos+1].cth, 1e-12))
      {
      if (ri[pos].cth > 0) p->r2 = ri[pos+1];
      else { p->r1 = ri[pos+1]; p->r2 = ri[pos]; }
      ++pos;
      }
    else
      {
      memset(&p->r2, 0, sizeof(p->r2));
      p->r2.nph = -1;
      }
    }
  free(ri);
  qsort(info->pair, (size_t)info->npairs, sizeof(orc_ringpair), cmp_pair);
  *geom_info = info;
  }

void orc_destroy_geom_info (orc_geom_info *info)
  { free(info->pair); free(info); }

/* HEALPix rings, libsharp/sharp_geomhelpers.c:38-93 */
void orc_make_weighted_healpix_geom_info (int nside, int stride,
  const double *weight, orc_geom_info **geom_info)
  {
  ptrdiff_t npix = (ptrdiff_t)nside*nside*12;
  ptrdiff_t ncap = 2*(ptrdiff_t)nside*(nside-1);
  int nrings = 4*nside-1;
  double *theta = xmalloc(sizeof(double)*nrings), *wgt = xmalloc(sizeof(double)*nrings),
         *phi0 = xmalloc(sizeof(double)*nrings);
  int *nph = xmalloc(sizeof(int)*nrings), *str = xmalloc(sizeof(int)*nrings);
  ptrdiff_t *ofs = xmalloc(sizeof(ptrdiff_t)*nrings);
  for (int m=0; m<nrings; ++m)
    {
    int ring = m+1;
    ptrdiff_t nr = (ring>2*nside) ? 4*nside-ring : ring;   /* northern index */
    str[m] = stride;
    if (nr < nside)
      {
      theta[m] = 2*asin(nr/(sqrt(6.)*nside));
      nph[m] = 4*(int)nr;
      phi0[m] = pi/nph[m];
      ofs[m] = 2*nr*(nr-1)*stride;
      }
    else
      {
      double fact1 = (8.*nside)/npix;
      theta[m] = acos((2*nside-nr)*fact1);
      nph[m] = 4*nside;
      phi0[m] = ((nr-nside)&1) ? 0 : pi/nph[m];
      ofs[m] = (ncap + (nr-nside)*nph[m])*stride;
      }
    if (nr != ring)
      {
      theta[m] = pi-theta[m];
      ofs[m] = (npix-nph[m])*stride - ofs[m];
      }
    wgt[m] = 4.*pi/npix*(weight ? weight[nr-1] : 1.);
    }
  orc_make_geom_info(nrings, nph, ofs, str, phi0, theta, wgt, geom_info);
  free(theta); free(wgt); free(phi0); free(nph); free(str); free(ofs);
  }

/* Gauss-Legendre nodes by Newton iteration, libsharp/sharp_geomhelpers.c:95-156 */
static double one_minus_x2 (double x)
  { return (fabs(x)>0.1) ? (1.+x)*(1.-x) : 1.-x*x; }

static void gauss_legendre (int n, double *x, double *w)
  {
  const double eps = 3e-14;
  int m = (n+1)>>1;
  double t0 = 1 - (1-1./n) / (8.*n*n);
  double t1 = 1./(4.*n+2.);
  for (int i=1; i<=m; ++i)
    {
    double x0 = cos(pi * ((i<<2)-1) * t1) * t0;
    int dobreak = 0, j = 0;
    double dpdx;
    for (;;)
      {
      double P_1 = 1.0, P0 = x0;
      for (int k=2; k<=n; k++)
        {
        double P_2 = P_1;
        P_1 = P0;
        P0 = x0*P_1 + (k-1.)/k * (x0*P_1-P_2);
        }
      dpdx = (P_1 - x0*P0) * n / one_minus_x2(x0);
      double x1 = x0 - P0/dpdx;
      double dx = x0-x1;
      x0 = x1;
      if (dobreak) break;
      if (fabs(dx)<=eps) dobreak = 1;
      if (++j>=100) fail("gauss_legendre", "convergence problem");
      }
    x[i-1] = -x0;
    x[n-i] = x0;
    w[i-1] = w[n-i] = 2. / (one_minus_x2(x0) * dpdx * dpdx);
    }
  }

static void make_iso_geom (int nrings, int nphi, double phi0, int stride_lon,
  int stride_lat, const double *theta, const double *weight,
  orc_geom_info **geom_info)
  {
  int *nph = xmalloc(sizeof(int)*nrings), *str = xmalloc(sizeof(int)*nrings);
  double *ph = xmalloc(sizeof(double)*nrings);
  ptrdiff_t *ofs = xmalloc(sizeof(ptrdiff_t)*nrings);
  for (int m=0; m<nrings; ++m)
    {
    nph[m] = nphi; ph[m] = phi0; str[m] = stride_lon;
    ofs[m] = (ptrdiff_t)m*stride_lat;
    }
  orc_make_geom_info(nrings, nph, ofs, str, ph, theta, weight, geom_info);
  free(nph); free(str); free(ph); free(ofs);
  }

void orc_make_gauss_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info)
  {
  double *theta = xmalloc(sizeof(double)*nrings), *w = xmalloc(sizeof(double)*nrings);
  gauss_legendre(nrings, theta, w);
  for (int m=0; m<nrings; ++m)
    {
    theta[m] = acos(-theta[m]);
    w[m] *= 2*pi/nphi;
    }
  make_iso_geom(nrings, nphi, phi0, stride_lon, stride_lat, theta, w, geom_info);
  free(theta); free(w);
  }

/* Fejer / Clenshaw-Curtis weights (Waldvogel 2006) as in
   libsharp/sharp_geomhelpers.c:192-334; the backward real FFT there is a
   direct DFT here. */
void orc_make_fejer1_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info)
  {
  double *theta = xmalloc(sizeof(double)*nrings), *w = xmalloc(sizeof(double)*nrings),
         *tmp = xmalloc(sizeof(double)*nrings);
  w[0] = 2.;
  for (int k=1; k<=(nrings-1)/2; ++k)
    {
    w[2*k-1] = 2./(1.-4.*k*k)*cos((k*pi)/nrings);
    w[2*k  ] = 2./(1.-4.*k*k)*sin((k*pi)/nrings);
    }
  if ((nrings&1)==0) w[nrings-1] = 0.;
  orc_rdft_backward(nrings, w, tmp);
  memcpy(w, tmp, sizeof(double)*nrings);
  for (int m=0; m<(nrings+1)/2; ++m)
    {
    theta[m] = pi*(m+0.5)/nrings;
    theta[nrings-1-m] = pi-theta[m];
    w[m] = w[nrings-1-m] = w[m]*2*pi/(nrings*(double)nphi);
    }
  make_iso_geom(nrings, nphi, phi0, stride_lon, stride_lat, theta, w, geom_info);
  free(theta); free(w); free(tmp);
  }

void orc_make_cc_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info)
  {
  int n = nrings-1;
  double *theta = xmalloc(sizeof(double)*nrings), *w = xmalloc(sizeof(double)*nrings),
         *tmp = xmalloc(sizeof(double)*nrings);
  for (int i=0; i<nrings; ++i) w[i] = 0.;
  double dw = -1./(n*n-1.+(n&1));
  w[0] = 2.+dw;
  for (int k=1; k<=(n/2-1); ++k)
    w[2*k-1] = 2./(1.-4.*k*k) + dw;
  w[2*(n/2)-1] = (n-3.)/(2*(n/2)-1) -1. -dw*((2-(n&1))*n-1);
  orc_rdft_backward(n, w, tmp);
  memcpy(w, tmp, sizeof(double)*n);
  w[n] = w[0];
  for (int m=0; m<(nrings+1)/2; ++m)
    {
    theta[m] = pi*m/(nrings-1.);
    if (theta[m]<1e-15) theta[m] = 1e-15;
    theta[nrings-1-m] = pi-theta[m];
    w[m] = w[nrings-1-m] = w[m]*2*pi/(n*(double)nphi);
    }
  make_iso_geom(nrings, nphi, phi0, stride_lon, stride_lat, theta, w, geom_info);
  free(theta); free(w); free(tmp);
  }

void orc_make_fejer2_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info)
  {
  int n = nrings+1;
  double *theta = xmalloc(sizeof(double)*nrings), *w = xmalloc(sizeof(double)*n),
         *tmp = xmalloc(sizeof(double)*n);
  for (int i=0; i<n; ++i) w[i] = 0.;
  w[0] = 2.;
  for (int k=1; k<=(n/2-1); ++k)
    w[2*k-1] = 2./(1.-4.*k*k);
  w[2*(n/2)-1] = (n-3.)/(2*(n/2)-1) -1.;
  orc_rdft_backward(n, w, tmp);
  for (int m=0; m<nrings; ++m) w[m] = tmp[m+1];
  for (int m=0; m<(nrings+1)/2; ++m)
    {
    theta[m] = pi*(m+1)/(nrings+1.);
    theta[nrings-1-m] = pi-theta[m];
    w[m] = w[nrings-1-m] = w[m]*2*pi/(n*(double)nphi);
    }
  make_iso_geom(nrings, nphi, phi0, stride_lon, stride_lat, theta, w, geom_info);
  free(theta); free(w); free(tmp);
  }

void orc_make_mw_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info)
  {
  double *theta = xmalloc(sizeof(double)*nrings);
  for (int m=0; m<nrings; ++m)
    {
    theta[m] = pi*(2.*m+1.)/(2.*nrings-1.);
    if (theta[m]>pi-1e-15) theta[m] = pi-1e-15;
    }
  make_iso_geom(nrings, nphi, phi0, stride_lon, stride_lat, theta, NULL, geom_info);
  free(theta);
  }

/* ------------------------------------------------------------------------
   a_lm layouts, libsharp/sharp.c:137-177 and sharp_almhelpers.c:35-70
   ------------------------------------------------------------------------ */
void orc_make_general_alm_info (int lmax, int nm, int stride, const int *mval,
  const ptrdiff_t *mvstart, int flags, orc_alm_info **alm_info)
  {
  orc_alm_info *info = xmalloc(sizeof(*info));
  info->lmax = lmax; info->nm = nm; info->stride = stride; info->flags = flags;
  info->mval = xmalloc(sizeof(int)*(size_t)nm);
  info->mvstart = xmalloc(sizeof(ptrdiff_t)*(size_t)nm);
  for (int i=0; i<nm; ++i) { info->mval[i] = mval[i]; info->mvstart[i] = mvstart[i]; }
  *alm_info = info;
  }

void orc_make_triangular_alm_info (int lmax, int mmax, int stride,
  orc_alm_info **alm_info)
  {
  int *mval = xmalloc(sizeof(int)*(mmax+1));
  ptrdiff_t *mvs = xmalloc(sizeof(ptrdiff_t)*(mmax+1));
  ptrdiff_t tval = 2*lmax+1;
  for (ptrdiff_t m=0; m<=mmax; ++m)
    { mval[m] = (int)m; mvs[m] = stride*((m*(tval-m))>>1); }
  orc_make_general_alm_info(lmax, mmax+1, stride, mval, mvs, 0, alm_info);
  free(mval); free(mvs);
  }

void orc_make_rectangular_alm_info (int lmax, int mmax, int stride,
  orc_alm_info **alm_info)
  {
  int *mval = xmalloc(sizeof(int)*(mmax+1));
  ptrdiff_t *mvs = xmalloc(sizeof(ptrdiff_t)*(mmax+1));
  for (ptrdiff_t m=0; m<=mmax; ++m)
    { mval[m] = (int)m; mvs[m] = stride*m*(lmax+1); }
  orc_make_general_alm_info(lmax, mmax+1, stride, mval, mvs, 0, alm_info);
  free(mval); free(mvs);
  }

void orc_destroy_alm_info (orc_alm_info *info)
  { free(info->mval); free(info->mvstart); free(info); }

/* ------------------------------------------------------------------------
   m cutoff per ring, libsharp/sharp.c:66-78
   ------------------------------------------------------------------------ */
int orc_get_mlim (int lmax, int spin, double sth, double cth)
  {
  double ofs = lmax*0.01;
  if (ofs<100.) ofs = 100.;
  double b = -2*spin*fabs(cth);
  double t1 = lmax*sth+ofs;
  double c = (double)spin*spin-t1*t1;
  double discr = b*b-4*c;
  if (discr<=0) return lmax;
  double res = (-b+sqrt(discr))/2.;
  if (res>lmax) res = lmax;
  return (int)(res+0.5);
  }

/* ------------------------------------------------------------------------
   exponent scaling, libsharp/sharp_core_inc.c:73-180 (one lane)
   A value v with scale s stands for v * 2^(800*s).
   ------------------------------------------------------------------------ */
static void normalize (double *val, int *scale, double maxval)
  {
  while (fabs(*val) > maxval) { *val *= fsmall; ++*scale; }
  if (*val != 0.)
    while (fabs(*val) < maxval*fsmall) { *val *= fbig; --*scale; }
  }

static void mypow (double val, int npow, double *resd, int *ress)
  {
  int scale = 0, scaleint = 0;
  double res = 1.;
  normalize(&val, &scaleint, fbighalf);
  do
    {
    if (npow&1)
      {
      res *= val;
      scale += scaleint;
      normalize(&res, &scale, fbighalf);
      }
    val *= val;
    scaleint += scaleint;
    normalize(&val, &scaleint, fbighalf);
    }
  while (npow >>= 1);
  *resd = res; *ress = scale;
  }

static int rescale (double *lam1, double *lam2, int *scale)
  {
  if (fabs(*lam2) > ftol)
    {
    *lam1 *= fsmall; *lam2 *= fsmall; ++*scale;
    return 1;
    }
  return 0;
  }

static double corfac (int scale)
  { return (scale<0) ? 0. : ((scale==0) ? 1. : fbig); }

/* ------------------------------------------------------------------------
   spin-0 Legendre functions, libsharp/sharp_ylmgen_c.c:44-78,145-161 (tables),
   sharp_core_inc.c:182-207 (start value) and sharp_core_inc2.c:164-220
   (scaled and unscaled loops).
   ------------------------------------------------------------------------ */
void orc_ylm_spin0 (int lmax, int m, double cth, double sth, double *out)
  {
  const double inv_sqrt4pi = 0.2820947917738781434740397257803862929220;
  for (int l=m; l<=lmax; ++l) out[l-m] = 0.;
  double mfac = inv_sqrt4pi;
  for (int k=1; k<=m; ++k) mfac *= sqrt((2*k+1.)/(2*k));
  /* rf[l] = {sqrt(2l+3)/sqrt((l+1+m)(l+1-m))*sqrt(2l+1),
              same * sqrt((l+m)(l-m))/sqrt(2l-1)}, rf[m].f1 = 0 */
  double (*rf)[2] = xmalloc(sizeof(double[2])*(size_t)(lmax+2));
  rf[m][0] = sqrt(2*m+3.); rf[m][1] = 0.;
  for (int l=m+1; l<=lmax; ++l)
    {
    double tmp = sqrt(2*l+3.)*(1./sqrt(l+1.+m))*(1./sqrt(l+1.-m));
    rf[l][0] = tmp*sqrt(2*l+1.);
    rf[l][1] = tmp*sqrt((double)(l+m))*sqrt((double)(l-m))*(1./sqrt(2*l-1.));
    }

  int l = m, scale;
  double lam1 = 0., lam2;
  mypow(sth, m, &lam2, &scale);
  lam2 *= (m&1) ? -mfac : mfac;
  normalize(&lam2, &scale, ftol);
  /* phase 1: no contribution while the value is below the IEEE limit */
  while (scale < 1)
    {
    if (l+2 > lmax) { free(rf); return; }
    lam1 = (cth*lam2)*rf[l][0] - lam1*rf[l][1];
    lam2 = (cth*lam1)*rf[l+1][0] - lam2*rf[l+1][1];
    rescale(&lam1, &lam2, &scale);
    l += 2;
    }
  /* phase 2: scaled recurrence, values multiplied by the correction factor */
  double cf = corfac(scale);
  while (scale < 0)
    {
    out[l-m] = lam2*cf;
    if (++l > lmax) { free(rf); return; }
    lam1 = (cth*lam2)*rf[l-1][0] - lam1*rf[l-1][1];
    out[l-m] = lam1*cf;
    if (++l > lmax) { free(rf); return; }
    lam2 = (cth*lam1)*rf[l-1][0] - lam2*rf[l-1][1];
    if (rescale(&lam1, &lam2, &scale)) cf = corfac(scale);
    }
  /* phase 3: plain IEEE recurrence */
  lam1 *= cf; lam2 *= cf;
  while (l < lmax)
    {
    lam1 = (cth*lam2)*rf[l][0] - lam1*rf[l][1];
    out[l-m] = lam2;
    out[l+1-m] = lam1;
    lam2 = (cth*lam1)*rf[l+1][0] - lam2*rf[l+1][1];
    l += 2;
    }
  if (l == lmax) out[l-m] = lam2;
  free(rf);
  }

/* ------------------------------------------------------------------------
   spin-s functions (Wigner-d based), libsharp/sharp_ylmgen_c.c:79-121,162-197
   (tables), sharp_core_inc.c:209-293 (start values, recurrence step) and
   sharp_core_inc2.c:432-519 (scaled/unscaled loops).
   outp/outm are indexed by l-mhi with mhi = max(m,spin).
   ------------------------------------------------------------------------ */
void orc_ylm_spin (int lmax, int m, int spin, double cth, double sth,
  double *outp, double *outm)
  {
  int s = spin;
  int mlo = (m<s) ? m : s, mhi = (m<s) ? s : m;
  for (int l=mhi; l<=lmax; ++l) outp[l-mhi] = outm[l-mhi] = 0.;
  if (mhi > lmax) return;

  /* prefactor sqrt((2 mhi)! / ((mhi+mlo)! (mhi-mlo)!)) with scaling */
  double *fac = xmalloc(sizeof(double)*(size_t)(2*lmax+1));
  int *facscale = xmalloc(sizeof(int)*(size_t)(2*lmax+1));
  fac[0] = 1; facscale[0] = 0;
  for (int k=1; k<2*lmax+1; ++k)
    {
    fac[k] = fac[k-1]*sqrt((double)k);
    facscale[k] = facscale[k-1];
    normalize(&fac[k], &facscale[k], fbighalf);
    }
  double prefac = fac[2*mhi]/fac[mhi+mlo];
  int fscale = facscale[2*mhi]-facscale[mhi+mlo];
  normalize(&prefac, &fscale, fbighalf);
  prefac /= fac[mhi-mlo];
  fscale -= facscale[mhi-mlo];
  normalize(&prefac, &fscale, fbighalf);
  free(fac); free(facscale);

  /* recurrence coefficients fx[l] for l=mhi+1..lmax */
  double (*fx)[3] = xmalloc(sizeof(double[3])*(size_t)(lmax+2));
  for (int l=0; l<lmax+2; ++l) fx[l][0] = fx[l][1] = fx[l][2] = 0.;
  for (int l=mhi; l<lmax; ++l)
    {
    double t = sqrt(1./(l+m+1.))*sqrt(1./(l-m+1.))*sqrt(1./(l+s+1.))*sqrt(1./(l-s+1.));
    double lt = 2*l+1, l1 = l+1;
    double invl = (l==0) ? 0. : 1./l;
    fx[l+1][0] = l1*lt*t;
    fx[l+1][1] = m*s*invl*(1./(l+1));
    t = sqrt((l+m)/(l+m+1.))*sqrt((l-m)/(l-m+1.))*sqrt((l+s)/(l+s+1.))*sqrt((l-s)/(l-s+1.));
    fx[l+1][2] = t*l1*invl;
    }

  int cosPow, sinPow, preMinus_p = 0, preMinus_m = 0;
  if (mhi == m)
    {
    cosPow = mhi+s; sinPow = mhi-s;
    preMinus_p = preMinus_m = ((mhi-s)&1);
    }
  else
    {
    cosPow = mhi+m; sinPow = mhi-m;
    preMinus_m = ((mhi+m)&1);
    }

  /* start values from cos(theta/2), sin(theta/2) powers */
  double cth2 = sqrt((1.+cth)*0.5), sth2 = sqrt((1.-cth)*0.5);
  if (cth2 < 1e-15) cth2 = 1e-15;
  if (sth2 < 1e-15) sth2 = 1e-15;
  if (sth < 0 && cth < 0) cth2 = -cth2;
  if (sth < 0 && cth > 0) sth2 = -sth2;
  double ccp, ssp, csp, scp;
  int ccps, ssps, csps, scps;
  mypow(cth2, cosPow, &ccp, &ccps); mypow(sth2, sinPow, &ssp, &ssps);
  mypow(cth2, sinPow, &csp, &csps); mypow(sth2, cosPow, &scp, &scps);

  double rec2p = prefac*ccp, rec2m = prefac*csp;
  int scalep = fscale+ccps, scalem = fscale+csps;
  normalize(&rec2m, &scalem, fbighalf);
  normalize(&rec2p, &scalep, fbighalf);
  rec2p *= ssp; scalep += ssps;
  rec2m *= scp; scalem += scps;
  if (preMinus_p) rec2p = -rec2p;
  if (preMinus_m) rec2m = -rec2m;
  if (s&1) rec2p = -rec2p;
  normalize(&rec2m, &scalem, ftol);
  normalize(&rec2p, &scalep, ftol);

  double rec1p = 0., rec1m = 0.;
  int l = mhi;
#define STEP(xp,xm,yp,ym,f) do { \
    xp = ((cth-f[1])*(f[0]*yp)) - f[2]*xp; \
    xm = ((cth+f[1])*(f[0]*ym)) - f[2]*xm; } while (0)

  while (scalep < 1 && scalem < 1)
    {
    if (l+2 > lmax) { free(fx); return; }
    STEP(rec1p, rec1m, rec2p, rec2m, fx[l+1]);
    STEP(rec2p, rec2m, rec1p, rec1m, fx[l+2]);
    rescale(&rec1p, &rec2p, &scalep);
    rescale(&rec1m, &rec2m, &scalem);
    l += 2;
    }
  double cfp = corfac(scalep), cfm = corfac(scalem);
  while (scalep < 0 || scalem < 0)
    {
    outp[l-mhi] = rec2p*cfp; outm[l-mhi] = rec2m*cfm;
    if (++l > lmax) { free(fx); return; }
    STEP(rec1p, rec1m, rec2p, rec2m, fx[l]);
    outp[l-mhi] = rec1p*cfp; outm[l-mhi] = rec1m*cfm;
    if (++l > lmax) { free(fx); return; }
    STEP(rec2p, rec2m, rec1p, rec1m, fx[l]);
    int a = rescale(&rec1p, &rec2p, &scalep);
    int b = rescale(&rec1m, &rec2m, &scalem);
    if (a|b) { cfp = corfac(scalep); cfm = corfac(scalem); }
    }
  rec1p *= cfp; rec2p *= cfp; rec1m *= cfm; rec2m *= cfm;
  while (l < lmax)
    {
    STEP(rec1p, rec1m, rec2p, rec2m, fx[l+1]);
    outp[l-mhi] = rec2p; outm[l-mhi] = rec2m;
    outp[l+1-mhi] = rec1p; outm[l+1-mhi] = rec1m;
    STEP(rec2p, rec2m, rec1p, rec1m, fx[l+2]);
    l += 2;
    }
  if (l == lmax) { outp[l-mhi] = rec2p; outm[l-mhi] = rec2m; }
#undef STEP
  free(fx);
  }

/* normalisation of spin transforms, libsharp/sharp_ylmgen_c.c:200-232 */
void orc_get_norm (int lmax, int spin, double *out)
  {
  if (spin==0) { for (int l=0; l<=lmax; ++l) out[l] = 1.; return; }
  double spinsign = -1.;
  if (spin&1) spinsign = -spinsign;
  for (int l=0; l<=lmax; ++l)
    out[l] = (l<spin) ? 0. : spinsign*0.5*sqrt((2*l+1)/(4*pi));
  }

void orc_get_d1norm (int lmax, double *out)
  {
  for (int l=0; l<=lmax; ++l)
    out[l] = (l<1) ? 0. : 0.5*sqrt(l*(l+1.)*(2*l+1.)/(4*pi));
  }

/* ------------------------------------------------------------------------
   direct real DFTs in fftpack storage (libfftpack/fftpack.c rfftf1/rfftb1)
   ------------------------------------------------------------------------ */
static void twiddles (int n, double *c, double *s)
  {
  for (int k=0; k<n; ++k) { c[k] = cos(2*pi*k/n); s[k] = sin(2*pi*k/n); }
  }

void orc_rdft_forward (int n, const double *in, double *out)
  {
  double *c = xmalloc(sizeof(double)*n), *s = xmalloc(sizeof(double)*n);
  twiddles(n, c, s);
  for (int k=0; k<=n/2; ++k)
    {
    double re = 0., im = 0.;
    for (int j=0; j<n; ++j)
      {
      int idx = (int)(((long long)k*j) % n);
      re += in[j]*c[idx];
      im -= in[j]*s[idx];
      }
    if (k==0) out[0] = re;
    else if (2*k==n) out[n-1] = re;
    else { out[2*k-1] = re; out[2*k] = im; }
    }
  free(c); free(s);
  }

void orc_rdft_backward (int n, const double *in, double *out)
  {
  double *c = xmalloc(sizeof(double)*n), *s = xmalloc(sizeof(double)*n);
  twiddles(n, c, s);
  for (int j=0; j<n; ++j)
    {
    double acc = in[0];
    for (int k=1; 2*k<n; ++k)
      {
      int idx = (int)(((long long)k*j) % n);
      acc += 2.*(in[2*k-1]*c[idx] - in[2*k]*s[idx]);
      }
    if ((n&1)==0) acc += in[n-1]*((j&1) ? -1. : 1.);
    out[j] = acc;
    }
  free(c); free(s);
  }

/* ------------------------------------------------------------------------
   the job
   ------------------------------------------------------------------------ */
typedef struct
  {
  int type, spin, nmaps, nalm, ntrans, flags, lmax, mmax;
  void **alm, **map;
  const orc_geom_info *g;
  const orc_alm_info *a;
  double *norm_l;
  /* phase[((pair*2+hemi)*(mmax+1)+m)*ncomp + comp], ncomp=ntrans*nmaps */
  cplx *phase;
  } job_t;

#define DP(job) ((job)->flags & ORC_DP)

static double map_get (const job_t *j, int i, ptrdiff_t idx)
  { return DP(j) ? ((double *)j->map[i])[idx] : ((float *)j->map[i])[idx]; }
static void map_add (const job_t *j, int i, ptrdiff_t idx, double v)
  {
  if (DP(j)) ((double *)j->map[i])[idx] += v;
  else ((float *)j->map[i])[idx] += (float)v;
  }
static cplx cmap_get (const job_t *j, int i, ptrdiff_t idx)
  {
  cplx r;
  if (DP(j)) { r.r = ((double *)j->map[i])[2*idx]; r.i = ((double *)j->map[i])[2*idx+1]; }
  else { r.r = ((float *)j->map[i])[2*idx]; r.i = ((float *)j->map[i])[2*idx+1]; }
  return r;
  }
static void cmap_add (const job_t *j, int i, ptrdiff_t idx, cplx v)
  {
  if (DP(j)) { ((double *)j->map[i])[2*idx] += v.r; ((double *)j->map[i])[2*idx+1] += v.i; }
  else { ((float *)j->map[i])[2*idx] += (float)v.r; ((float *)j->map[i])[2*idx+1] += (float)v.i; }
  }

/* a_lm element addressing, libsharp/sharp.c:481-606: returns the real-unit
   offset; for PACKED m=0 entries only a real part exists. */
static ptrdiff_t alm_ofs (const orc_alm_info *a, int mi, int l, int *realonly)
  {
  int packed = a->flags & ORC_PACKED;
  ptrdiff_t ofs = a->mvstart[mi], stride = a->stride;
  if (!packed) ofs *= 2;
  *realonly = packed && (a->mval[mi]==0);
  if (!*realonly) stride *= 2;
  return ofs + l*stride;
  }
static cplx alm_get (const job_t *j, int i, int mi, int l)
  {
  int ro; ptrdiff_t o = alm_ofs(j->a, mi, l, &ro);
  cplx r;
  if (DP(j)) { r.r = ((double *)j->alm[i])[o]; r.i = ro ? 0. : ((double *)j->alm[i])[o+1]; }
  else { r.r = ((float *)j->alm[i])[o]; r.i = ro ? 0. : ((float *)j->alm[i])[o+1]; }
  return r;
  }
static void alm_add (const job_t *j, int i, int mi, int l, cplx v, int realonly_write)
  {
  int ro; ptrdiff_t o = alm_ofs(j->a, mi, l, &ro);
  if (DP(j))
    {
    ((double *)j->alm[i])[o] += v.r;
    if (!realonly_write) ((double *)j->alm[i])[o+1] += v.i;
    }
  else
    {
    ((float *)j->alm[i])[o] += (float)v.r;
    if (!realonly_write) ((float *)j->alm[i])[o+1] += (float)v.i;
    }
  }
static void alm_clear (const job_t *j, int i, int mi, int l)
  {
  int ro; ptrdiff_t o = alm_ofs(j->a, mi, l, &ro);
  if (DP(j)) { ((double *)j->alm[i])[o] = 0.; if (!ro) ((double *)j->alm[i])[o+1] = 0.; }
  else { ((float *)j->alm[i])[o] = 0.f; if (!ro) ((float *)j->alm[i])[o+1] = 0.f; }
  }

static const orc_ringinfo *ring_of (const job_t *j, int p, int h)
  { return h ? &j->g->pair[p].r2 : &j->g->pair[p].r1; }

/* libsharp/sharp.c:355-453 */
static void clear_outputs (job_t *j)
  {
  if (j->flags & ORC_ADD) return;
  if (j->type == ORC_MAP2ALM)
    {
    for (int i=0; i<j->ntrans*j->nalm; ++i)
      for (int mi=0; mi<j->a->nm; ++mi)
        for (int l=j->a->mval[mi]; l<=j->lmax; ++l)
          alm_clear(j, i, mi, l);
    }
  else
    {
    for (int i=0; i<j->ntrans*j->nmaps; ++i)
      for (int p=0; p<j->g->npairs; ++p)
        for (int h=0; h<2; ++h)
          {
          const orc_ringinfo *r = ring_of(j, p, h);
          for (ptrdiff_t k=0; k<r->nph; ++k)
            {
            ptrdiff_t idx = r->ofs + k*r->stride;
            if (j->flags & ORC_NO_FFT)
              {
              if (DP(j)) ((double *)j->map[i])[2*idx] = ((double *)j->map[i])[2*idx+1] = 0.;
              else ((float *)j->map[i])[2*idx] = ((float *)j->map[i])[2*idx+1] = 0.f;
              }
            else
              {
              if (DP(j)) ((double *)j->map[i])[idx] = 0.;
              else ((float *)j->map[i])[idx] = 0.f;
              }
            }
          }
    }
  }

static cplx cmul (cplx a, cplx b)
  { cplx r = { a.r*b.r - a.i*b.i, a.r*b.i + a.i*b.r }; return r; }

/* map -> phases for one ring: ringhelper_ring2phase (sharp.c:305-353) with a
   direct DFT, or ring2phase_direct (sharp.c:631-652) for NO_FFT. */
static void ring_to_phase (const job_t *j, const orc_ringinfo *r, cplx *ph,
  int comp)
  {
  int mmax = j->mmax, ncomp = j->ntrans*j->nmaps;
  if (r->nph <= 0)
    {
    for (int m=0; m<=mmax; ++m) ph[m*ncomp].r = ph[m*ncomp].i = 0.;
    return;
    }
  if (j->flags & ORC_NO_FFT)
    {
    if (r->nph != mmax+1) fail("ring_to_phase", "bad ring size");
    double wgt = (j->flags & ORC_USE_WEIGHTS) ? r->nph*r->weight : 1.;
    if (j->flags & ORC_REAL_HARMONICS) wgt *= sqrt_two;
    for (int m=0; m<=mmax; ++m)
      {
      cplx v = cmap_get(j, comp, r->ofs + (ptrdiff_t)m*r->stride);
      ph[m*ncomp].r = v.r*wgt; ph[m*ncomp].i = v.i*wgt;
      }
    return;
    }
  int n = r->nph;
  double wgt = (j->flags & ORC_USE_WEIGHTS) ? r->weight : 1.;
  if (j->flags & ORC_REAL_HARMONICS) wgt *= sqrt_two;
  double *c = xmalloc(sizeof(double)*n), *s = xmalloc(sizeof(double)*n),
         *f = xmalloc(sizeof(double)*n);
  twiddles(n, c, s);
  for (int k=0; k<n; ++k) f[k] = map_get(j, comp, r->ofs + (ptrdiff_t)k*r->stride);
  int norot = fabs(r->phi0) < 1e-14;
  for (int m=0; m<=mmax; ++m)
    {
    /* sum_j f_j exp(-2 pi i m j / n), i.e. bin (m mod n) of the forward FFT */
    cplx acc = {0., 0.};
    for (int k=0; k<n; ++k)
      {
      int idx = (int)(((long long)m*k) % n);
      acc.r += f[k]*c[idx];
      acc.i -= f[k]*s[idx];
      }
    acc.r *= wgt; acc.i *= wgt;
    if (!norot)
      {
      cplx sh = { cos(m*(-r->phi0)), sin(m*(-r->phi0)) };
      acc = cmul(acc, sh);
      }
    ph[m*ncomp] = acc;
    }
  free(c); free(s); free(f);
  }

/* phases -> map for one ring: ringhelper_phase2ring (sharp.c:253-303) with a
   direct DFT, or phase2ring_direct (sharp.c:653-669) for NO_FFT.  The
   aliasing fold of the reference is the identity
       f_j = Re c_0 + 2 sum_{m>=1} Re(c_m exp(i m phi_j)). */
static void phase_to_ring (const job_t *j, const orc_ringinfo *r,
  const cplx *ph, int comp)
  {
  int mmax = j->mmax, ncomp = j->ntrans*j->nmaps;
  if (r->nph <= 0) return;
  if (j->flags & ORC_NO_FFT)
    {
    if (r->nph != mmax+1) fail("phase_to_ring", "bad ring size");
    double wgt = (j->flags & ORC_USE_WEIGHTS) ? r->nph*r->weight : 1.;
    if (j->flags & ORC_REAL_HARMONICS) wgt *= sqrt_one_half;
    for (int m=0; m<=mmax; ++m)
      {
      cplx v = { ph[m*ncomp].r*wgt, ph[m*ncomp].i*wgt };
      cmap_add(j, comp, r->ofs + (ptrdiff_t)m*r->stride, v);
      }
    return;
    }
  int n = r->nph;
  double wgt = (j->flags & ORC_USE_WEIGHTS) ? r->weight : 1.;
  if (j->flags & ORC_REAL_HARMONICS) wgt *= sqrt_one_half;
  int norot = fabs(r->phi0) < 1e-14;
  cplx *cm = xmalloc(sizeof(cplx)*(mmax+1));
  for (int m=0; m<=mmax; ++m)
    {
    cplx v = { ph[m*ncomp].r*wgt, ph[m*ncomp].i*wgt };
    if (!norot && m>0)
      {
      cplx sh = { cos(m*r->phi0), sin(m*r->phi0) };
      v = cmul(v, sh);
      }
    cm[m] = v;
    }
  double *c = xmalloc(sizeof(double)*n), *s = xmalloc(sizeof(double)*n);
  twiddles(n, c, s);
  for (int k=0; k<n; ++k)
    {
    double acc = cm[0].r;
    for (int m=1; m<=mmax; ++m)
      {
      int idx = (int)(((long long)m*k) % n);
      acc += 2.*(cm[m].r*c[idx] - cm[m].i*s[idx]);
      }
    map_add(j, comp, r->ofs + (ptrdiff_t)k*r->stride, acc);
    }
  free(cm); free(c); free(s);
  }

/* synthesis for one (m, pair): sharp_core_inc2.c:634-724 */
static void synth_pair (job_t *j, int mi, int p, double *lp, double *lm,
  const cplx *almtmp)
  {
  int m = j->a->mval[mi], lmax = j->lmax, ncomp = j->ntrans*j->nmaps;
  const orc_ringpair *pr = &j->g->pair[p];
  cplx *phN = &j->phase[((size_t)(2*p+0)*(j->mmax+1)+m)*ncomp];
  cplx *phS = &j->phase[((size_t)(2*p+1)*(j->mmax+1)+m)*ncomp];
  int ispair = pr->r2.nph > 0;
  for (int c=0; c<ncomp; ++c) { phN[c].r = phN[c].i = phS[c].r = phS[c].i = 0.; }
  if (orc_get_mlim(lmax, j->spin, pr->r1.sth, pr->r1.cth) < m) return;
  if (j->spin == 0)
    {
    orc_ylm_spin0(lmax, m, pr->r1.cth, pr->r1.sth, lp);
    for (int t=0; t<j->ntrans; ++t)
      {
      cplx p1 = {0,0}, p2 = {0,0};
      for (int l=m; l<=lmax; ++l)
        {
        cplx a = almtmp[(size_t)(l-m)*j->ntrans + t];
        double lam = lp[l-m];
        if (((l-m)&1)==0) { p1.r += lam*a.r; p1.i += lam*a.i; }
        else { p2.r += lam*a.r; p2.i += lam*a.i; }
        }
      phN[t].r = p1.r+p2.r; phN[t].i = p1.i+p2.i;
      if (ispair) { phS[t].r = p1.r-p2.r; phS[t].i = p1.i-p2.i; }
      }
    return;
    }
  int s = j->spin, mhi = (m<s) ? s : m;
  orc_ylm_spin(lmax, m, s, pr->r1.cth, pr->r1.sth, lp, lm);
  int flip = (mhi-m+s)&1;
  for (int t=0; t<j->ntrans; ++t)
    {
    /* q/u accumulators for the two parity classes (p1: l-mhi even) */
    double q[2][2] = {{0,0},{0,0}}, u[2][2] = {{0,0},{0,0}};
    for (int l=mhi; l<=lmax; ++l)
      {
      cplx ag = almtmp[((size_t)(l-m)*j->ntrans + t)*j->nalm];
      cplx ac = {0., 0.};
      if (j->nalm == 2) ac = almtmp[((size_t)(l-m)*j->ntrans + t)*2 + 1];
      double rp = lp[l-mhi], rm = lm[l-mhi];
      double lw = rp+rm, lx = rm-rp;
      int x = (l-mhi)&1, y = x^1;
      q[x][0] += ag.r*lw; q[x][1] += ag.i*lw;
      u[x][0] += ac.r*lw; u[x][1] += ac.i*lw;
      q[y][0] -= ac.i*lx; q[y][1] += ac.r*lx;
      u[y][0] += ag.i*lx; u[y][1] -= ag.r*lx;
      }
    cplx *QN = &phN[2*t], *UN = &phN[2*t+1], *QS = &phS[2*t], *US = &phS[2*t+1];
    QN->r = q[0][0]+q[1][0]; QN->i = q[0][1]+q[1][1];
    UN->r = u[0][0]+u[1][0]; UN->i = u[0][1]+u[1][1];
    if (ispair)
      {
      double sg = flip ? -1. : 1.;
      QS->r = sg*(q[0][0]-q[1][0]); QS->i = sg*(q[0][1]-q[1][1]);
      US->r = sg*(u[0][0]-u[1][0]); US->i = sg*(u[0][1]-u[1][1]);
      }
    }
  }

/* analysis for one (m, pair): sharp_core_inc2.c:725-794 */
static void anal_pair (job_t *j, int mi, int p, double *lp, double *lm,
  cplx *almtmp)
  {
  int m = j->a->mval[mi], lmax = j->lmax, ncomp = j->ntrans*j->nmaps;
  const orc_ringpair *pr = &j->g->pair[p];
  const cplx *phN = &j->phase[((size_t)(2*p+0)*(j->mmax+1)+m)*ncomp];
  const cplx *phS = &j->phase[((size_t)(2*p+1)*(j->mmax+1)+m)*ncomp];
  int ispair = pr->r2.nph > 0;
  if (orc_get_mlim(lmax, j->spin, pr->r1.sth, pr->r1.cth) < m) return;
  if (j->spin == 0)
    {
    orc_ylm_spin0(lmax, m, pr->r1.cth, pr->r1.sth, lp);
    for (int t=0; t<j->ntrans; ++t)
      {
      cplx n = phN[t], s = ispair ? phS[t] : (cplx){0., 0.};
      cplx p1 = { n.r+s.r, n.i+s.i }, p2 = { n.r-s.r, n.i-s.i };
      for (int l=m; l<=lmax; ++l)
        {
        cplx *a = &almtmp[(size_t)(l-m)*j->ntrans + t];
        const cplx *pp = ((l-m)&1) ? &p2 : &p1;
        a->r += lp[l-m]*pp->r; a->i += lp[l-m]*pp->i;
        }
      }
    return;
    }
  int s = j->spin, mhi = (m<s) ? s : m;
  orc_ylm_spin(lmax, m, s, pr->r1.cth, pr->r1.sth, lp, lm);
  int flip = (mhi-m+s)&1;
  for (int t=0; t<j->ntrans; ++t)
    {
    cplx nq = phN[2*t], nu = phN[2*t+1], sq = {0,0}, su = {0,0};
    if (ispair) { sq = phS[2*t]; su = phS[2*t+1]; }
    if (flip) { sq.r = -sq.r; sq.i = -sq.i; su.r = -su.r; su.i = -su.i; }
    /* P[0] = N+S (used with l-mhi even), P[1] = N-S; each {qr,qi,ur,ui} */
    double P[2][4] = {
      { nq.r+sq.r, nq.i+sq.i, nu.r+su.r, nu.i+su.i },
      { nq.r-sq.r, nq.i-sq.i, nu.r-su.r, nu.i-su.i } };
    for (int l=mhi; l<=lmax; ++l)
      {
      double rp = lp[l-mhi], rm = lm[l-mhi];
      double lw = rp+rm, lx = rm-rp;
      int x = (l-mhi)&1, y = x^1;
      cplx *ag = &almtmp[((size_t)(l-m)*j->ntrans + t)*2];
      cplx *ac = ag+1;
      ag->r += P[x][0]*lw - P[y][3]*lx;
      ag->i += P[x][1]*lw + P[y][2]*lx;
      ac->r += P[x][2]*lw + P[y][1]*lx;
      ac->i += P[x][3]*lw - P[y][0]*lx;
      }
    }
  }

void orc_execute (int type, int spin, void *alm, void *map,
  const orc_geom_info *geom_info, const orc_alm_info *alm_info, int ntrans,
  int flags, double *time, unsigned long long *opcnt)
  {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  job_t j;
  /* job normalisation, libsharp/sharp.c:840-867 */
  if ((ntrans<=0) || (ntrans>100)) fail("orc_execute", "bad number of simultaneous transforms");
  if (type==ORC_DERIV1) spin = 1;
  if (type==ORC_MAP2ALM) flags |= ORC_USE_WEIGHTS;
  if (type==ORC_Yt) type = ORC_MAP2ALM;
  if (type==ORC_WY) { type = ORC_ALM2MAP; flags |= ORC_USE_WEIGHTS; }
  if ((spin<0) || (spin>alm_info->lmax)) fail("orc_execute", "bad spin");
  j.type = type; j.spin = spin; j.flags = flags; j.ntrans = ntrans;
  j.nmaps = (type==ORC_DERIV1) ? 2 : ((spin>0) ? 2 : 1);
  j.nalm = (type==ORC_DERIV1) ? 1 : ((spin>0) ? 2 : 1);
  j.alm = (void **)alm; j.map = (void **)map;
  j.g = geom_info; j.a = alm_info;
  j.lmax = alm_info->lmax;
  /* all m from 0..nm-1 exactly once, libsharp/sharp.c:238-251 */
  {
  int nm = alm_info->nm;
  int *chk = calloc((size_t)nm+1, sizeof(int));
  for (int i=0; i<nm; ++i)
    {
    int mc = alm_info->mval[i];
    if ((mc<0) || (mc>=nm)) fail("orc_execute", "not all m values are present");
    if (chk[mc]) fail("orc_execute", "duplicate m value");
    chk[mc] = 1;
    }
  free(chk);
  j.mmax = nm-1;
  }
  if (j.lmax < j.mmax) fail("orc_execute", "incorrect l_max: must be >= m_max");
  j.norm_l = xmalloc(sizeof(double)*(j.lmax+1));
  if (type==ORC_DERIV1) orc_get_d1norm(j.lmax, j.norm_l);
  else orc_get_norm(j.lmax, spin, j.norm_l);

  clear_outputs(&j);

  int npairs = geom_info->npairs, ncomp = ntrans*j.nmaps;
  size_t nph_tot = (size_t)npairs*2*(j.mmax+1)*ncomp;
  j.phase = xmalloc(sizeof(cplx)*nph_tot);
  memset(j.phase, 0, sizeof(cplx)*nph_tot);

  double *lp = xmalloc(sizeof(double)*(j.lmax+2)), *lm = xmalloc(sizeof(double)*(j.lmax+2));
  int na = ntrans*j.nalm;
  cplx *almtmp = xmalloc(sizeof(cplx)*(size_t)(j.lmax+1)*na);
  unsigned long long triples = 0;

  if (type == ORC_MAP2ALM)
    {
    for (int p=0; p<npairs; ++p)
      for (int h=0; h<2; ++h)
        for (int c=0; c<ncomp; ++c)
          ring_to_phase(&j, ring_of(&j, p, h),
            &j.phase[((size_t)(2*p+h)*(j.mmax+1))*ncomp + c], c);
    }

  for (int mi=0; mi<alm_info->nm; ++mi)
    {
    int m = alm_info->mval[mi];
    double n0 = 1.;
    /* a_lm -> almtmp (libsharp/sharp.c:481-545) */
    if (type != ORC_MAP2ALM)
      {
      if (flags & ORC_REAL_HARMONICS) n0 = sqrt_two;
      for (int l=m; l<=j.lmax; ++l)
        for (int i=0; i<na; ++i)
          {
          cplx a = alm_get(&j, i, mi, l);
          if (m==0) a.i = 0.;
          double f = j.norm_l[l]*((m==0) ? n0 : 1.);
          almtmp[(size_t)(l-m)*na+i].r = a.r*f;
          almtmp[(size_t)(l-m)*na+i].i = a.i*f;
          }
      }
    else
      memset(almtmp, 0, sizeof(cplx)*(size_t)(j.lmax+1)*na);

    for (int p=0; p<npairs; ++p)
      {
      const orc_ringpair *pr = &geom_info->pair[p];
      if (orc_get_mlim(j.lmax, spin, pr->r1.sth, pr->r1.cth) >= m)
        triples += (unsigned long long)(j.lmax-m+1);
      if (type == ORC_MAP2ALM) anal_pair(&j, mi, p, lp, lm, almtmp);
      else synth_pair(&j, mi, p, lp, lm, almtmp);
      }

    /* almtmp -> a_lm (libsharp/sharp.c:547-606) */
    if (type == ORC_MAP2ALM)
      {
      if (flags & ORC_REAL_HARMONICS) n0 = sqrt_one_half;
      for (int l=m; l<=j.lmax; ++l)
        for (int i=0; i<na; ++i)
          {
          cplx x = almtmp[(size_t)(l-m)*na+i];
          double f = j.norm_l[l]*((m==0) ? n0 : 1.);
          x.r *= f; x.i *= f;
          alm_add(&j, i, mi, l, x, m==0);
          }
      }
    }

  if (type != ORC_MAP2ALM)
    {
    for (int p=0; p<npairs; ++p)
      for (int h=0; h<2; ++h)
        for (int c=0; c<ncomp; ++c)
          phase_to_ring(&j, ring_of(&j, p, h),
            &j.phase[((size_t)(2*p+h)*(j.mmax+1))*ncomp + c], c);
    }

  free(lp); free(lm); free(almtmp); free(j.phase); free(j.norm_l);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (time) *time = (t1.tv_sec-t0.tv_sec) + 1e-9*(t1.tv_nsec-t0.tv_nsec);
  if (opcnt)
    {
    unsigned long long per = (type==ORC_DERIV1) ? 12+8ull*ntrans :
      ((spin==0) ? 4+4ull*ntrans : 12+16ull*ntrans);
    *opcnt = triples*per;
    }
  }

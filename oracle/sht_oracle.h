/*
 * sht_oracle.h -- CPU restatement of libsharp's sharp_execute path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP engine: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only to compute expected values.  The product library never links,
 * loads or calls it.
 *
 * Scope: everything sharp_execute does for one call, written as scalar C
 * (one "vector lane" per ring pair), with the ring transforms evaluated as
 * direct DFTs instead of FFTs.  It is meant to be read as the specification
 * of the path and to finish in seconds at the small sizes used by the tests
 * (nside <= 64, lmax <= 256).
 *
 * The info structs below have the same memory layout as the reference's
 * sharp_ringinfo / sharp_ringpair / sharp_geom_info / sharp_alm_info
 * (libsharp/sharp_lowlevel.h:43-93), so the same objects can be handed to
 * the oracle, to the reference build and to the drop-in library.
 */
#ifndef SHT_ORACLE_H
#define SHT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
  {
  double theta, phi0, weight, cth, sth;
  ptrdiff_t ofs;
  int nph, stride;
  } orc_ringinfo;

typedef struct { orc_ringinfo r1, r2; } orc_ringpair;

typedef struct
  {
  orc_ringpair *pair;
  int npairs, nphmax;
  } orc_geom_info;

typedef struct
  {
  int lmax;
  int nm;
  int *mval;
  int flags;
  ptrdiff_t *mvstart;
  ptrdiff_t stride;
  } orc_alm_info;

/* job types and flags: same numeric values as sharp_jobtype / sharp_jobflags
   (libsharp/sharp_lowlevel.h:162-191) */
enum { ORC_MAP2ALM=0, ORC_ALM2MAP=1, ORC_Yt=2, ORC_WY=3, ORC_DERIV1=4 };
enum { ORC_DP=1<<4, ORC_ADD=1<<5, ORC_REAL_HARMONICS=1<<6, ORC_NO_FFT=1<<7,
       ORC_USE_WEIGHTS=1<<20, ORC_PACKED=1 };

/* --- info objects ------------------------------------------------------ */
void orc_make_geom_info (int nrings, const int *nph, const ptrdiff_t *ofs,
  const int *stride, const double *phi0, const double *theta,
  const double *wgt, orc_geom_info **geom_info);
void orc_make_weighted_healpix_geom_info (int nside, int stride,
  const double *weight, orc_geom_info **geom_info);
void orc_make_gauss_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info);
void orc_make_fejer1_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info);
void orc_make_cc_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info);
void orc_make_fejer2_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info);
void orc_make_mw_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, orc_geom_info **geom_info);
void orc_destroy_geom_info (orc_geom_info *info);

void orc_make_triangular_alm_info (int lmax, int mmax, int stride,
  orc_alm_info **alm_info);
void orc_make_rectangular_alm_info (int lmax, int mmax, int stride,
  orc_alm_info **alm_info);
void orc_make_general_alm_info (int lmax, int nm, int stride, const int *mval,
  const ptrdiff_t *mvstart, int flags, orc_alm_info **alm_info);
void orc_destroy_alm_info (orc_alm_info *info);

/* --- building blocks, exported so the tests can pin them one by one ------ */
int orc_get_mlim (int lmax, int spin, double sth, double cth);
/* lambda_lm(theta) for l=m..lmax (spin 0) with the reference's recurrence and
   exponent scaling; out[l-m].  Values whose scale is below the IEEE range are
   returned as 0, exactly as the accumulation loops drop them. */
void orc_ylm_spin0 (int lmax, int m, double cth, double sth, double *out);
/* spin-s recurrence pair (lambda_plus, lambda_minus) for l=max(m,s)..lmax */
void orc_ylm_spin (int lmax, int m, int spin, double cth, double sth,
  double *outp, double *outm);
void orc_get_norm (int lmax, int spin, double *out);
void orc_get_d1norm (int lmax, double *out);
/* direct real DFTs with fftpack conventions (libfftpack/ls_fft.h:119-130):
   forward: r0 r1 i1 r2 i2 ... ; backward: inverse of that, unnormalised */
void orc_rdft_forward (int n, const double *in, double *out);
void orc_rdft_backward (int n, const double *in, double *out);

/* --- the transform ------------------------------------------------------- */
void orc_execute (int type, int spin, void *alm, void *map,
  const orc_geom_info *geom_info, const orc_alm_info *alm_info, int ntrans,
  int flags, double *time, unsigned long long *opcnt);

#ifdef __cplusplus
}
#endif

#endif

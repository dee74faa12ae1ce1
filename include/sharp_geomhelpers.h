/*
 * sharp_geomhelpers.h -- standard map geometries.
 * Same declarations as libsharp/sharp_geomhelpers.h:46-135.
 */
#ifndef PLANCK_SHARP_GEOMHELPERS_H
#define PLANCK_SHARP_GEOMHELPERS_H

#include "sharp_lowlevel.h"

#ifdef __cplusplus
extern "C" {
#endif

/* HEALPix map of the given nside; weight has 2*nside entries (relative ring
   weights, NULL = all 1).  libsharp/sharp_geomhelpers.h:46-47 */
void sharp_make_weighted_healpix_geom_info (int nside, int stride,
  const double *weight, sharp_geom_info **geom_info);

/* libsharp/sharp_geomhelpers.h:52-54 */
static inline void sharp_make_healpix_geom_info (int nside, int stride,
  sharp_geom_info **geom_info)
  { sharp_make_weighted_healpix_geom_info (nside, stride, NULL, geom_info); }

/* Gauss-Legendre grid.  libsharp/sharp_geomhelpers.h:63-64 */
void sharp_make_gauss_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, sharp_geom_info **geom_info);

/* Equidistant grid without poles (Fejer's first rule).
   libsharp/sharp_geomhelpers.h:79-80 */
void sharp_make_fejer1_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, sharp_geom_info **geom_info);

/* Former name of the Fejer-1 grid.  libsharp/sharp_geomhelpers.h:84-89 */
static inline void sharp_make_ecp_geom_info (int nrings, int nphi, double phi0,
  int stride_lon, int stride_lat, sharp_geom_info **geom_info)
  {
  sharp_make_fejer1_geom_info (nrings, nphi, phi0, stride_lon, stride_lat,
    geom_info);
  }

/* Equidistant grid with poles (Clenshaw-Curtis).
   libsharp/sharp_geomhelpers.h:103-104 */
void sharp_make_cc_geom_info (int nrings, int ppring, double phi0,
  int stride_lon, int stride_lat, sharp_geom_info **geom_info);

/* Equidistant grid, Fejer's second rule.  libsharp/sharp_geomhelpers.h:118-119 */
void sharp_make_fejer2_geom_info (int nrings, int ppring, double phi0,
  int stride_lon, int stride_lat, sharp_geom_info **geom_info);

/* McEwen & Wiaux grid, no quadrature weights.
   libsharp/sharp_geomhelpers.h:134-135 */
void sharp_make_mw_geom_info (int nrings, int ppring, double phi0,
  int stride_lon, int stride_lat, sharp_geom_info **geom_info);

#ifdef __cplusplus
}
#endif

#endif

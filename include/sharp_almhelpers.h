/*
 * sharp_almhelpers.h -- standard a_lm storage schemes.
 * Same declarations as libsharp/sharp_almhelpers.h:44-51.
 */
#ifndef PLANCK_SHARP_ALMHELPERS_H
#define PLANCK_SHARP_ALMHELPERS_H

#include "sharp_lowlevel.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Healpix_cxx layout: index = m*(2*lmax+1-m)/2 + l (times stride).
   libsharp/sharp_almhelpers.h:44-45 */
void sharp_make_triangular_alm_info (int lmax, int mmax, int stride,
  sharp_alm_info **alm_info);

/* Fortran Healpix layout: index = m*(lmax+1) + l (times stride).
   libsharp/sharp_almhelpers.h:50-51 */
void sharp_make_rectangular_alm_info (int lmax, int mmax, int stride,
  sharp_alm_info **alm_info);

#ifdef __cplusplus
}
#endif

#endif

/*
 * sharp_lowlevel.h -- C interface of the MI355X spherical harmonic transform
 * engine.  Drop-in for libsharp's low-level interface:
 *   libsharp/sharp_lowlevel.h:43-93   (types; identical memory layout)
 *   libsharp/sharp_lowlevel.h:95-152  (a_lm and geometry info objects)
 *   libsharp/sharp_lowlevel.h:160-227 (job types, flags, sharp_execute)
 *
 * Every function keeps the reference's name, argument meaning and error
 * behaviour: invalid arguments print "file, line (func):\nmessage" to stderr
 * and terminate the process with exit status 1.
 *
 * sharp_execute takes host pointers.  The arrays are staged to the GPU, the
 * transform runs there, and the results are copied back before it returns.
 * Device-resident entry points are declared in sharp_hip.h.
 */
#ifndef PLANCK_SHARP_LOWLEVEL_H
#define PLANCK_SHARP_LOWLEVEL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One iso-latitude ring (libsharp/sharp_lowlevel.h:43-48). */
typedef struct
  {
  double theta, phi0, weight, cth, sth;
  ptrdiff_t ofs;
  int nph, stride;
  } sharp_ringinfo;

/* Two rings placed symmetrically about the equator; r2.nph<0 when the ring
   has no partner (libsharp/sharp_lowlevel.h:52-55). */
typedef struct
  {
  sharp_ringinfo r1,r2;
  } sharp_ringpair;

/* Map geometry (libsharp/sharp_lowlevel.h:59-63). */
typedef struct
  {
  sharp_ringpair *pair;
  int npairs, nphmax;
  } sharp_geom_info;

/* a_lm storage scheme (libsharp/sharp_lowlevel.h:70-85):
   index(l, mval[mi]) = mvstart[mi] + stride*l. */
typedef struct
  {
  int lmax;
  int nm;
  int *mval;
  int flags;
  ptrdiff_t *mvstart;
  ptrdiff_t stride;
  } sharp_alm_info;

/* libsharp/sharp_lowlevel.h:88-93: with SHARP_PACKED the m=0 coefficients
   carry no imaginary part; mvstart is in real units for all m, stride in
   real units for m=0 and complex units otherwise. */
typedef enum { SHARP_PACKED = 1 } sharp_almflags;

/* libsharp/sharp_lowlevel.h:106-107 */
void sharp_make_alm_info (int lmax, int mmax, int stride,
  const ptrdiff_t *mstart, sharp_alm_info **alm_info);
/* libsharp/sharp_lowlevel.h:119-120 */
void sharp_make_general_alm_info (int lmax, int nm, int stride, const int *mval,
  const ptrdiff_t *mvstart, int flags, sharp_alm_info **alm_info);
/* libsharp/sharp_lowlevel.h:125 */
ptrdiff_t sharp_alm_index (const sharp_alm_info *self, int l, int mi);
/* libsharp/sharp_lowlevel.h:127 */
void sharp_destroy_alm_info (sharp_alm_info *info);

/* libsharp/sharp_lowlevel.h:148-150: rings are paired north/south and the
   pairs sorted by (nph, phi0, cos(theta)), exactly as libsharp does. */
void sharp_make_geom_info (int nrings, const int *nph, const ptrdiff_t *ofs,
  const int *stride, const double *phi0, const double *theta,
  const double *wgt, sharp_geom_info **geom_info);
/* libsharp/sharp_lowlevel.h:153 */
void sharp_destroy_geom_info (sharp_geom_info *info);

/* Job types (libsharp/sharp_lowlevel.h:162-169). */
typedef enum { SHARP_YtW=0,
               SHARP_MAP2ALM=SHARP_YtW,
               SHARP_Y=1,
               SHARP_ALM2MAP=SHARP_Y,
               SHARP_Yt=2,
               SHARP_WY=3,
               SHARP_ALM2MAP_DERIV1=4
             } sharp_jobtype;

/* Job flags (libsharp/sharp_lowlevel.h:172-191).  The low four bits (nv)
   select a CPU vector blocking in libsharp; they are accepted and ignored. */
typedef enum { SHARP_DP              = 1<<4,
               SHARP_ADD             = 1<<5,
               SHARP_REAL_HARMONICS  = 1<<6,
               SHARP_NO_FFT          = 1<<7,
               SHARP_USE_WEIGHTS     = 1<<20,
               SHARP_NO_OPENMP       = 1<<21,
               SHARP_NVMAX           = (1<<4)-1
             } sharp_jobflags;

/* libsharp/sharp_lowlevel.h:222-224.  alm and map are arrays of host
   pointers (ntrans*nalm and ntrans*nmaps entries).  *time receives the wall
   time of the call, *opcnt the operation count of libsharp's flop model
   (libsharp/sharp_core_inc2.c:171-173,440-442,583-585) summed over all
   (m, ring pair, l) with m <= sharp_get_mlim() of the pair. */
void sharp_execute (sharp_jobtype type, int spin, void *alm, void *map,
  const sharp_geom_info *geom_info, const sharp_alm_info *alm_info, int ntrans,
  int flags, double *time, unsigned long long *opcnt);

/* libsharp/sharp_lowlevel.h:226-227.  Kept for ABI compatibility: the GPU
   engine does not chunk ring pairs by these values but by device memory. */
void sharp_set_chunksize_min(int new_chunksize_min);
void sharp_set_nchunks_max(int new_nchunks_max);

/* libsharp/sharp_internal.h:62-64 (exported by libsharp, used by its tests) */
int sharp_get_nv_max (void);
int sharp_nv_oracle (sharp_jobtype type, int spin, int ntrans);
int sharp_get_mlim (int lmax, int spin, double sth, double cth);

#ifdef __cplusplus
}
#endif

#endif

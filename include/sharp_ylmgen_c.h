/*
 * sharp_ylmgen_c.h -- helper data of the Y_lm recurrences (host side).
 * Same types and declarations as libsharp/sharp_ylmgen_c.h:36-95: the
 * generator struct has libsharp's memory layout field for field, and the
 * tables it holds after sharp_Ylmgen_prepare have libsharp's values (the GPU
 * engine builds its device tables with the same per-m routines).
 */
#ifndef PLANCK_SHARP_YLMGEN_C_H
#define PLANCK_SHARP_YLMGEN_C_H

#ifdef __cplusplus
extern "C" {
#endif

/* exponent-scaling range (libsharp/sharp_ylmgen_c.h:39): a value v with
   scale s stands for v * 2^(800 s) */
enum { sharp_minscale=0, sharp_limscale=1, sharp_maxscale=1 };
static const double sharp_fbig=0x1p+800, sharp_fsmall=0x1p-800;
static const double sharp_ftol=0x1p-60;
static const double sharp_fbighalf=0x1p+400;

typedef struct { double f[2]; } sharp_ylmgen_dbl2;
typedef struct { double f[3]; } sharp_ylmgen_dbl3;

/* libsharp/sharp_ylmgen_c.h:47-73 */
typedef struct
  {
  int lmax, mmax, s;
  double *cf;        /* [maxscale-minscale+1]: 2^(800 (i+minscale)) */
  int m;             /* the m of the last prepare (-1 before any) */
  /* spin 0 */
  double *mfac;      /* [mmax+1] lambda_mm prefactors */
  sharp_ylmgen_dbl2 *rf;   /* [lmax+1] recurrence coefficients of m */
  /* spin s */
  int sinPow, cosPow, preMinus_p, preMinus_m;
  double *prefac;    /* [mmax+1] start-value prefactors */
  int *fscale;       /* [mmax+1] their scales */
  sharp_ylmgen_dbl3 *fx;   /* [lmax+2] recurrence coefficients of m */
  /* internal: spin 0 */
  double *root, *iroot;
  /* internal: spin s */
  double *flm1, *flm2, *inv;
  int mlo, mhi;
  } sharp_Ylmgen_C;

/* libsharp/sharp_ylmgen_c.h:77: tables up to l=l_max, m=m_max.  Errors as
   in libsharp: "incorrect spin: must be nonnegative", "incorrect l_max: must
   be >= spin", "incorrect l_max: must be >= m_max" (message, exit(1)). */
void sharp_Ylmgen_init (sharp_Ylmgen_C *gen, int l_max, int m_max, int spin);

/* libsharp/sharp_ylmgen_c.h:80 */
void sharp_Ylmgen_destroy (sharp_Ylmgen_C *gen);

/* libsharp/sharp_ylmgen_c.h:83: fills rf (spin 0) or fx, cosPow, sinPow,
   preMinus_p/m (spin s) for this m ("incorrect m" for m < 0). */
void sharp_Ylmgen_prepare (sharp_Ylmgen_C *gen, int m);

/* libsharp/sharp_ylmgen_c.h:88: norm_l of a spin (lmax+1 values, caller
   frees with free()). */
double *sharp_Ylmgen_get_norm (int lmax, int spin);

/* libsharp/sharp_ylmgen_c.h:94: norm_l of the first-derivative maps. */
double *sharp_Ylmgen_get_d1norm (int lmax);

#ifdef __cplusplus
}
#endif

#endif

/*
 * sharp_hip.h -- entry points of the MI355X engine beyond libsharp's
 * interface.  None of these exists in the reference; they serve callers whose
 * data already lives in GPU memory (the reference equivalent of each one is
 * named in its comment).
 *
 * Conventions: `alm` and `map` are host arrays of DEVICE pointers with the
 * same meaning and count as in sharp_execute (libsharp/sharp_lowlevel.h:
 * 196-224); `stream` is a hipStream_t (NULL: the null stream of the current
 * device, as in HIP).  The call enqueues the work and returns; the caller
 * synchronises the stream before reading the outputs.  Each stream has its
 * own library work buffers (kept until sharp_hip_release), so calls on
 * different streams may run concurrently.  Argument errors behave as in
 * sharp_execute (message, then exit(1)).
 */
#ifndef SHARP_HIP_H
#define SHARP_HIP_H

#include "sharp_lowlevel.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Precision of the Legendre stage, or-ed into `flags` (a bit libsharp does
   not use).  Default: FP64 throughout.  The mixed form keeps the lambda_lm
   recurrence, its start values and its exponent scaling in FP64 and runs the
   [rings x l] . [l x maps] contraction on the FP32 matrix cores with FP32
   accumulation; its outputs stay within 1e-6 relative L2 of the FP64 form. */
#define SHARP_HIP_MIXED_FP32    (1<<24)
/* As SHARP_HIP_MIXED_FP32, with the contraction on the BF16 matrix cores:
   every FP32 operand is split exactly into three BF16 pieces and the six
   piece products of order >= 2^-16 are accumulated in FP32 (the dropped ones
   are below 2^-20 of each product, about 2^-24 in the rms); outputs stay
   within 1e-6 relative L2 of the FP64 form. */
#define SHARP_HIP_MIXED_BF16X3  (1<<25)

/* Device-resident form of sharp_execute (libsharp/sharp_lowlevel.h:222-224).
   *opcnt as in sharp_execute; no timing is taken. */
void sharp_hip_execute_device (sharp_jobtype type, int spin, void *alm,
  void *map, const sharp_geom_info *geom_info, const sharp_alm_info *alm_info,
  int ntrans, int flags, void *stream, unsigned long long *opcnt);

/* Adjoint of synthesis, x -> Y^T x: map2alm without quadrature weights
   (sharp_execute with SHARP_Yt, libsharp/sharp.c:848). */
void sharp_hip_alm2map_adjoint_device (int spin, void *alm, void *map,
  const sharp_geom_info *geom_info, const sharp_alm_info *alm_info,
  int ntrans, int flags, void *stream);

/* Adjoint of analysis, a -> W Y a: alm2map with quadrature weights
   (sharp_execute with SHARP_WY, libsharp/sharp.c:849). */
void sharp_hip_map2alm_adjoint_device (int spin, void *alm, void *map,
  const sharp_geom_info *geom_info, const sharp_alm_info *alm_info,
  int ntrans, int flags, void *stream);

/* sharp_hip_execute_device with a per-l factor on the a_lm side: synthesis
   computes map = Y (f_l a_lm) without modifying the caller's a_lm, analysis
   a_lm (+)= f_l (Y^T W map).  Replaces Alm::ScaleL before alm2map or after
   map2alm, e.g. smoothWithGauss (Healpix_cxx/alm_powspec_tools.cc:250-283)
   followed by alm2map in smoothing_cxx (Healpix_cxx/smoothing_cxx_module.cc:
   79-81), with the factor applied while the a_lm are gathered or scattered.
   lfac: host array of lmax+1 doubles, read before the call returns. */
void sharp_hip_execute_scaled_device (sharp_jobtype type, int spin, void *alm,
  void *map, const sharp_geom_info *geom_info, const sharp_alm_info *alm_info,
  int ntrans, int flags, const double *lfac, void *stream);

/* Jacobi-iterated analysis, map2alm_iter / map2alm_pol_iter
   (Healpix_cxx/alm_healpix_tools.cc:73-85, 239-265) on device arrays:
   a = Y^T W m, then num_iter times a += Y^T W (m - Y a).  One spin per call
   (T and E/B of a polarised map are independent).  The residual maps live in
   a library work buffer; SHARP_ADD is rejected. */
void sharp_hip_map2alm_iter_device (int spin, void *alm, void *map,
  const sharp_geom_info *geom_info, const sharp_alm_info *alm_info,
  int ntrans, int flags, int num_iter, void *stream);

/* Batches of transforms of uneven size (several geometries, lmax, spins and
   transform counts), no counterpart in the reference (whose callers issue one
   sharp_execute per geometry).  `alm`/`map` are host arrays of device
   pointers as in sharp_execute.  All jobs of a batch are syntheses
   (SHARP_ALM2MAP, SHARP_WY, SHARP_ALM2MAP_DERIV1) or all analyses
   (SHARP_MAP2ALM, SHARP_Yt), with the same `flags`: FP64, or one of the
   mixed forms (SHARP_HIP_MIXED_FP32, SHARP_HIP_MIXED_BF16X3).  The Legendre
   stage of all jobs runs as one persistent launch per kernel kind (spin 0,
   spin s, deriv1) over a work list of tiles of every job, dealt to the XCDs
   by cost.  The tile shape -- transforms per tile, dispatch (persistent
   workgroups per CU taking tiles from per-XCD counters, or, for syntheses,
   one flat launch), orders m per synthesis workgroup (8 or 4, one per
   wavefront of 64 ring pairs) or ring pairs per analysis workgroup (256 or
   128) -- is chosen by timing the batch at creation
   (cached by batch shape; SHARP_HIP_TUNE=0 disables, SHARP_HIP_TUNE_CACHE
   names a file that keeps the choices).  With use_graph the whole batch is
   recorded into a HIP graph at creation and each execute replays it: the
   device buffers of the jobs must stay the same (their contents may
   change).  Executes of one batch must not overlap in time (they share its
   work buffers and counters); different batches may run concurrently. */
typedef struct {
  int type;                         /* a sharp_jobtype */
  int spin;
  void *alm;                        /* host array of device pointers */
  void *map;                        /* host array of device pointers */
  const sharp_geom_info *geom_info;
  const sharp_alm_info *alm_info;
  int ntrans;
} sharp_hip_job;

typedef struct sharp_hip_batch sharp_hip_batch;

sharp_hip_batch *sharp_hip_batch_create (int njobs, const sharp_hip_job *jobs,
  int flags, int use_graph);
/* Enqueues the batch on `stream` (NULL: the null stream). */
void sharp_hip_batch_execute (sharp_hip_batch *batch, void *stream);
/* The tuned configuration: transforms per tile, persistent workgroups per
   CU (0: the synthesis runs as one flat launch with a wavefront per tile),
   and whether a graph is replayed. */
void sharp_hip_batch_config (const sharp_hip_batch *batch, int *tgroup,
  int *wg_per_cu, int *graph);
/* The tuned tile shape: orders m per synthesis workgroup (0 for analysis
   batches) and wavefronts of 64 ring pairs per analysis workgroup (0 for
   synthesis batches). */
void sharp_hip_batch_tiles (const sharp_hip_batch *batch, int *mblock,
  int *ring_waves);
void sharp_hip_batch_destroy (sharp_hip_batch *batch);

/* Builds (and caches) the device tables for this geometry / a_lm layout /
   spin outside any timed region.  Optional: the first transform does it.
   Also finds the current device's scratch area of
   sharp_hip_execute_device_noalloc. */
void sharp_hip_prepare (const sharp_geom_info *geom_info,
  const sharp_alm_info *alm_info, int spin);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): sharp_hip_execute_device with the same arguments and
   results, which calls no hipMalloc, hipFree or other allocating routine.
   Its work memory (phase array, canonical a_lm, analysis partial sums, BF16
   pieces) comes from a fixed scratch area of SHARP_HIP_NOALLOC_SCRATCH_BYTES
   per GPU, a device array each GPU holds from the first sharp_hip_prepare on
   it until the process ends.  Requirements, each checked with a clear error
   (message, then exit(1)):
   * sharp_hip_prepare(geom_info, alm_info, spin) was called on this device
     (spin 1 for SHARP_ALM2MAP_DERIV1) since the last sharp_hip_release;
   * the job's work memory fits the scratch area (about Nside 256, lmax 512,
     TQU, one transform; the message states the bytes needed).
   All calls on one GPU share its scratch area, which keeps the last call's
   contents: calls that overlap in time on one GPU (different streams,
   threads) race, so serialise them.  Per-stage profiling skips these calls. */
#define SHARP_HIP_NOALLOC_SCRATCH_BYTES (64u << 20)
void sharp_hip_execute_device_noalloc (sharp_jobtype type, int spin,
  void *alm, void *map, const sharp_geom_info *geom_info,
  const sharp_alm_info *alm_info, int ntrans, int flags, void *stream,
  unsigned long long *opcnt);

/* Waits for the work enqueued on `stream` (NULL: the null stream). */
void sharp_hip_synchronize (void *stream);

/* Waits for the device, then releases every cached table and work buffer of
   the current device.  Live batches keep the tables their jobs use (they
   are freed with the batch), so a batch created before stays usable. */
void sharp_hip_release (void);

/* Per-stage timing (the reference reports only whole-call time and a flop
   count, libsharp/sharp_lowlevel.h:218-221).  sharp_hip_profile(1) clears the
   counters and records a pair of HIP events on the launch stream around each
   stage of every later call; sharp_hip_profile(0) stops.
   sharp_hip_profile_stats waits for the recorded events and fills, for up to
   maxn stages: names (32 chars each), total milliseconds, launches, and the
   algorithmic flops (libsharp's flop model) and compulsory HBM bytes of those
   launches.  Returns the number of stages. */
void sharp_hip_profile (int enable);
int sharp_hip_profile_stats (int maxn, char *names, double *ms,
  long long *launches, double *flops, double *bytes);

/* Number of visible GPUs (0 on a host without one). */
int sharp_hip_device_count (void);

#ifdef __cplusplus
}
#endif

#endif

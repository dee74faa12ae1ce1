// kernels.hpp -- launch interface between the host engine and the HIP
// kernels.  Plain structs of device pointers; no torch types.
#pragma once

#include <hip/hip_runtime.h>
#include "shp_internal.hpp"
#include "sharp_hip.h"

namespace shp {

// Arguments of the Legendre (m <-> ring pair) stage.
struct LegArgs {
  int lmax, mmax, nslots, nmb;
  int spin;             // 0, or s>0 for the Wigner-d recurrences
  int deriv1;           // 1: spin-1 synthesis of a gradient (no curl a_lm)
  int ncomp_alm;        // components in almtmp   (ntrans*nalm)
  int ncomp_map;        // components in phase    (ntrans*nmaps)
  int t0;               // first transform handled by this launch
  const double *cth, *sth;
  const int *mlim, *ispair;
  const double *tab;    // recurrence coefficients, rows at tab_ofs(m,lmax)
  int unit_tab;         // 1: tab holds the coefficients of lambda_l / s_l whose
                        // lambda_{l-2} factor is 1 (spin 0: b_l; spin s: f2);
                        // the mixed-precision kernels need it
  const double *mfac;   // spin 0: lambda_mm prefactors [mmax+1]
  const double *prefac; // spin s: start-value prefactors [mmax+1]
  const int *fscale;    // spin s: their scales [mmax+1]
  const double2 *alm;   // synthesis input  (canonical almtmp)
  double2 *almout;      // analysis output  (canonical almtmp)
  double2 *phase;
  // mixed-precision (FP32 MFMA) forms only
  int ntl;              // transforms in this launch
  int ngroups;          // analysis: ring-pair groups per m (partial sums)
  const int *gmax;      // analysis: largest mlim of each group [ngroups]
  void *part;           // analysis: partial a_lm [ngroups][rows][part_ld], float
                        // (mixed form) or double (FP64 matrix-core form)
  int64_t part_ld;      // floats per partial row
  int part_col0;        // float column of this launch's first transform
  // analysis launches cover the orders m0 <= m < m1 (m0 a multiple of 8);
  // the partial buffer holds their rows: canonical row r of group g at
  // (g * part_rows + r - part_row0) * part_ld
  int m0, m1;
  int64_t part_row0, part_rows;
  // BF16 spin-s synthesis: the a_lm split into BF16 pieces by the gather
  // (bs_pairs), else nullptr
  const unsigned *bsplit;
  const int64_t *bs_base;
  int bs_ncol;
};

// BF16 pieces of the FP32 a_lm of a BF16 spin-s synthesis, written by the
// a_lm gather as the B operand words of the synthesis's first accumulator.
// Output column c = 4t+k is (Q.re, Q.im, U.re, U.im)[k] of transform t; its
// lambda_w entry is X_c = (E.re, E.im, B.re, B.im)[k] and its lambda_x entry
// Y_c = s_k X_(c^3), s_k = -1 for k = 0, 3 and +1 else (libsharp/
// sharp_core_inc2.c:282-306; without B, as for DERIV1, X = (E.re, E.im, 0, 0)).
// For each m, column c (4*ntrans of them, then one zero column) and piece q
// (h, m, l of the exact three-piece split) the rows l = o, o+1, ... from
// o = max(m, spin) form pairs r: one 32-bit word holds X_c of row o+2r (low
// half) and Y_c of row o+2r+1 (high half), zero past lmax; bs_pairs words per
// (c, q), starting at word bs_base[m] + (3c + q) * bs_pairs.  The second
// accumulator's word (Y_c of row o+2r, X_c of row o+2r+1) is s_k times the
// word of column c^3.
__host__ __device__ inline int bs_pairs(int lmax, int o)
{
  const int p = (lmax - o + 2) / 2;
  return p > 0 ? p : 0;
}

// One tile of the persistent batched Legendre kernels: job index, transform
// group (first transform, count, partial-sum column) and the tile's
// coordinates (synthesis: m-block, 64-slot block; analysis: m, slot group).
struct BatchItem {
  int job, t0, ntl, col0;
  int major, minor, pad0, pad1;
};
enum { kBatchSynth0, kBatchSynthS, kBatchSynthD, kBatchAnal0, kBatchAnalS };
// One launch of a batched Legendre kernel kind over its work list.
struct BatchLaunch {
  int kind;
  int ncb;          // 16-column blocks of the mixed forms (1 or 2)
  int contr;        // kContr*
  int shape;        // wavefronts per workgroup: synthesis 4 or 8 (items are
                    // (m-block of `shape` orders m, slot block) tiles, one order
                    // m per wavefront), analysis 2 or 4 (items are (m, ring group
                    // of 64*shape slots) tiles, whose largest m limits are in the
                    // job's gmax)
  int tg;           // transforms of every item (FP64 syntheses), else 0
  bool flat;        // syntheses: one item per workgroup, grid flat_grid;
                    // else persistent workgroups taking items from counters
  int grid;         // persistent workgroups (a multiple of 8; capped at what
                    // stays resident)
  int flat_grid;    // flat workgroups: 8 x the largest part
  const BatchItem *items;
  const int *xofs;  // [9]: the work list's parts for the 8 XCDs
  int *ctr;         // [8]: the parts' item counters (zeroed by the launch)
};
void launch_batch(const BatchLaunch &l, const LegArgs *jobs, hipStream_t st);
// the FP64 syntheses of a batch (legendre.hip): every item of the list has
// tg transforms, 1, 2 or 4 (spin 0) / 1 or 2 (spin s)
void launch_batch_synth_f64(int kind, int tg, int shape, bool flat, const LegArgs *jobs,
                            const BatchItem *items, const int *xofs, int *ctr, int grid,
                            int flat_grid, hipStream_t st);
// the same items on the FP64 matrix cores (synth64 tiles, ncb blocks of 16
// real columns)
void launch_batch_synth64(int kind, int ncb, int shape, bool flat, const LegArgs *jobs,
                          const BatchItem *items, const int *xofs, int *ctr, int grid,
                          int flat_grid, hipStream_t st);

// ntg = transforms in this launch (1, 2 or 4 for spin 0; 1 or 2 for spin s)
void launch_synth(const LegArgs &a, int ntg, hipStream_t st);
void launch_anal(const LegArgs &a, int ntg, hipStream_t st);
// FP64 synthesis with the contraction on the FP64 matrix cores: a.ntl
// transforms (up to kF64SynT0 for spin 0, kF64SynTS for spin s) per launch
constexpr int kF64SynT0 = 16;
constexpr int kF64SynTS = 8;
void launch_synth64(const LegArgs &a, hipStream_t st);

// Mixed-precision forms: FP64 recurrence, FP32 MFMA contraction over a batch
// of up to kMixMaxT transforms per launch (a.ntl).
constexpr int kMixMaxT0 = 16;   // spin 0: 2 real columns per transform
constexpr int kMixMaxTS = 8;    // spin s: 4 real columns per transform
// analysis: wavefronts per workgroup (spin 0 / spin s); a workgroup sums the
// ring pairs of one group of 64*waves slots
#ifndef SHP_MIX_WAVES0
#define SHP_MIX_WAVES0 4
#endif
#ifndef SHP_MIX_WAVESS
#define SHP_MIX_WAVESS 4
#endif
constexpr int kMixWaves0 = SHP_MIX_WAVES0;
constexpr int kMixWavesS = SHP_MIX_WAVESS;
constexpr int mix_group_slots(int spin) { return 64 * (spin == 0 ? kMixWaves0 : kMixWavesS); }
// contraction of the matrix-core Legendre kernels: FP32 MFMA; FP64 MFMA
// (the FP64 form: FP64 partial sums, up to 8 spin-0 / 8 spin-s transforms
// per launch); BF16 MFMA on exact three-piece BF16 splits of FP32 operands
enum { kContrF32 = 0, kContrF64 = 1, kContrBf16x3 = 2 };
void launch_msynth(const LegArgs &a, int contr, hipStream_t st);
void launch_manal(const LegArgs &a, int contr, hipStream_t st);
constexpr int kF64MaxT0 = 8;
constexpr int kF64MaxTS = 8;

// Arguments of the ring stage (phase <-> map).
constexpr int kMaxComp = 200;   // ntrans <= 100, two components each

// Per-component device pointers, passed by value in the kernel arguments.
struct PtrTab {
  void *p[kMaxComp];
};

struct RingArgs {
  int nslots, nmb, mmax, ncomp;
  int flags;            // SHARP_* flags after job normalisation
  int64_t map_base;     // subtracted from every pixel index
  const RingDev *ring;  // [nslots][2]
  double2 *phase;
  int nphmax;
  int dp;               // 1: double maps, 0: float maps
  int max_lds;          // bytes of LDS a workgroup may allocate
  PtrTab maps;          // [ncomp] device pointers (float or double)
};
// Direct-summation ring transforms over the rings in `list` (nlist entries of
// slot*2+hemi); exact for any ring length, O(nph*mmax) per ring.  Used for
// SHARP_NO_FFT and for rings too long for the LDS FFT.
void launch_phase2ring_dft(const RingArgs &a, const int *list, int nlist, hipStream_t st);
void launch_ring2phase_dft(const RingArgs &a, const int *list, int nlist, hipStream_t st);

// Per-ring FFT plan.  rad[] lists the Stockham radices of N (smooth lengths)
// or of the Bluestein length M (8, then 8s and one 4 or 2, then a last 8; at
// M = 8192: 8, 16, 8, 8).
struct RingFftPlan {
  int ring;        // slot*2 + hemi
  int n;           // pixels
  int N;           // complex FFT length: n/2 (even n) or n (odd n)
  int bluestein;   // 1: N has a prime factor > 7
  int M;           // Bluestein convolution length (power of two, >= 2N-1, >= 64)
  int nrad;
  int8_t rad[16];
  int64_t boff;    // Bluestein: offset of the plan's block in the table
                   // (kernel-spectrum entries 0..M/2, then chirp entries
                   // 0..N/2: the halves of two symmetric sequences)
};
constexpr int kNumFftClasses = 5;   // LDS buffers of 512 << c complex doubles
// kinds of plan in a class: bluestein + 2 * radix set (0: 8, 4, 2, 3, 5, 7;
// 1: 8, 4, 2 only; 2: 8 only)
constexpr int kFftKinds = 6;
// plans[kFftKinds*class + kind], sorted by decreasing ring length; hplans
// are host copies of the same arrays
struct FftClasses {
  const RingFftPlan *plans[kFftKinds * kNumFftClasses];
  const RingFftPlan *hplans[kFftKinds * kNumFftClasses];
  int count[kFftKinds * kNumFftClasses];
  const double2 *btab;   // Bluestein chirp spectra
};
// Fills btab at the offsets of the given plans (one plan per distinct (N, M)).
void launch_bluestein_table(const FftClasses &uniq, double2 *btab, hipStream_t st);
// writes the ring FFTs' table of roots of unity on the current device (once)
void ensure_twiddles(hipStream_t st);
// Two more streams of the launch stream's device and the events that fork
// them from it and join them back: with them the ring FFT classes of one
// stage run on three streams, so that a class's last workgroups overlap the
// next classes' first ones instead of leaving CUs idle.
struct FftFork {
  hipStream_t aux[2];
  hipEvent_t fork, join[2];
};
void launch_ring_fft(bool synth, const RingArgs &a, const FftClasses &fc, hipStream_t st,
                     const FftFork *fk = nullptr);

// Rings too long for every LDS class: one complex FFT of the ring's length n
// through global memory (power-of-two n), or Bluestein with L = 2^lg >= 2n-1.
struct LongRingPlan {
  int ring;        // slot*2 + hemi (for the Bluestein table: unused)
  int n;           // pixels
  int lg;          // log2 of the FFT length L
  int bluestein;   // 1: n is not a power of two
  int64_t boff;    // Bluestein: offset of the L spectrum entries in the table
};
struct LongRingArgs {
  const LongRingPlan *plans;
  int nplans;
  const double2 *btab;   // Bluestein kernel spectra
  double2 *scratch;      // nslots_wg slots of `slot` complex values
  int64_t slot;          // 2 * the largest L
};
// nslots_wg workgroups, each with its scratch slot, take the (ring,
// component) items in turn
void launch_long_rings(bool synth, const RingArgs &a, const LongRingArgs &la, int nslots_wg,
                       hipStream_t st);
// fills btab at each plan's boff (plans with bluestein set, one per n)
void launch_long_table(const LongRingPlan *plans, int nplans, double2 *scratch, int64_t slot,
                       int nslots_wg, double2 *btab, hipStream_t st);

// a_lm in the caller's layout <-> canonical almtmp.
struct AlmArgs {
  int lmax, mmax, ncomp;
  int flags;            // job flags (ADD, REAL_HARMONICS, DP)
  int packed;           // alm_info flags & SHARP_PACKED
  int64_t stride;       // alm_info stride
  int64_t alm_base;     // subtracted from every a_lm index (real units)
  const int64_t *mvstart_by_m;  // [mmax+1]
  const double *norm_l; // [lmax+1]
  // with a unit table (the mixed-precision forms): s_{m,l} at
  // tri_ofs(m,lmax)+l-m, applied beside norm_l; else nullptr
  const double *lsc;
  double2 *almtmp;
  int tmp_f32;          // gather: almtmp holds float2 (the mixed syntheses)
  PtrTab alms;          // [ncomp] device pointers (float or double data)
  // analysis from partial sums (mixed forms): when part != nullptr the
  // components [c0, c0+npc) are sum_g part[g] over the groups g with
  // gmax[g] >= m, instead of almtmp; component c of a row is at float
  // 2*(c - part_c0) of it, rows are part_ld floats apart
  const void *part;
  int part_f64;         // 1: part holds doubles
  const int *gmax;
  int ngroups, c0, npc;
  int64_t part_ld;
  int part_c0;
  // the scatter writes the orders m0 <= m < m1; partial rows as in LegArgs
  int m0, m1;
  int64_t part_row0, part_rows;
  // BF16 spin-s synthesis: where launch_alm_gather_split puts the BF16
  // pieces (bs_pairs), else nullptr
  unsigned *bsplit;
  const int64_t *bs_base;
  int bs_ncol, spin;
  int bs_deriv;         // the a_lm are E only (DERIV1): X = (E.re, E.im, 0, 0)
};
// dst[c][k] = src[c][k] - dst[c][k] for k < n, c < ncomp (float or double)
void launch_map_residual(const PtrTab &src, const PtrTab &dst, int ncomp, int64_t n,
                         bool dp, hipStream_t st);
void launch_alm_gather(const AlmArgs &a, hipStream_t st);
void launch_alm_scatter(const AlmArgs &a, hipStream_t st);
// the gather of a BF16 spin-s synthesis: FP32 rows and their BF16 pieces
// (bs_pairs)
void launch_alm_gather_split(const AlmArgs &a, hipStream_t st);

// noalloc.hip: the fixed scratch area of sharp_hip_execute_device_noalloc,
// one copy per GPU.  noalloc_setup() finds the current device's copy (once
// per device); noalloc_scratch(dev) returns it, or nullptr before set-up.
// Both are called with the engine's mutex held.
constexpr size_t kNoallocScratchBytes = SHARP_HIP_NOALLOC_SCRATCH_BYTES;
void noalloc_setup();
void *noalloc_scratch(int device);
// base[m] = word offset of m's BF16 pieces (the layout of bs_pairs), for
// m = 0..mmax
void launch_bs_base(int64_t *base, int lmax, int mmax, int spin, int ncol, hipStream_t st);

} // namespace shp

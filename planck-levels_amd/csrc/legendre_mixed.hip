// legendre_mixed.hip -- mixed-precision Legendre stage on CDNA4 (gfx950).
//
// The lambda_lm recurrence, its start values and its exponent scaling run in
// FP64 exactly as in the FP64 kernels (same operations, same order, same
// switch points between the scaled and the unscaled domain).  Only the
// contraction of lambda with the batch of transforms runs on the FP32 matrix
// cores, v_mfma_f32_16x16x4f32:  C[16x16] += A[16x4] . B[4x16], where lane
// i+16k holds A[i][k] and B[k][i], and lane n+16q holds C[4q..4q+3][n].
//
// Synthesis: one wavefront per (m, 64 ring pairs), lane = ring pair.  A lane
// produces lambda for 8 consecutive l (spin 0) or lambda_w / lambda_x for 4
// consecutive l (spin s); per accumulator that is one K=4 slice of A with the
// ring pair in the lane and the l index in the register.  Two permlane32 and
// two permlane16 swaps per 4 registers exchange the 16-lane rows so that the
// register index becomes the lane row, which is A's K layout.  B is the
// [4 l x 16 columns] slice of the a_lm, read from the canonical FP64 rows and
// rounded to FP32.  The N/S parity sums p1/p2 of the FP64 kernel are separate
// MFMA accumulators; for spin s the lambda_x terms go to the accumulator of
// the other parity with the rotated a_lm columns (libsharp/sharp_core_inc2.c:
// 282-306).  The MFMA accumulators are emptied into FP32 totals every 8 K
// slices, which keeps the rounding error of long sums near unit roundoff.
//
// Analysis: a workgroup of kMixWaves0/S wavefronts per (m, group of 64*waves
// ring pairs).  Each lane runs the recurrence for its ring pair 32 l at a time
// and stores lambda (FP32) in its wavefront's LDS slice; the MFMA reads it back
// transposed, A = [16 l of one parity x 4 ring pairs], against
// B = [4 ring pairs x 16 columns], the phases of the wavefront's 64 ring pairs,
// which stay in registers for the whole l range.  The wavefronts' [16 l x 16]
// tiles are added in LDS in a fixed order and stored as the group's partial
// a_lm (FP32); the a_lm scatter adds the groups in FP64 in a fixed order, so
// results do not depend on scheduling.
//
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2;
// xcd_map keeps all workgroups of one m-block (synthesis) on one XCD, and
// xcd_map_mblock those of the 8 m of one phase block (analysis), so the
// recurrence tables, a_lm rows and phase lines of those m are fetched into
// one L2 only.

#include <hip/hip_runtime.h>
#include <type_traits>
#include "kernels.hpp"
#include "legendre_common.hpp"

namespace shp {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline f4 mfma4(float a, float b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

typedef double d4 __attribute__((ext_vector_type(4)));

// 16x16x4 matrix products in FP32 or FP64.  A and B share one lane layout;
// the accumulator rows differ: lane n+16q, register v holds row 4q+v in
// FP32 and row 4v+q in FP64.
template <typename T> struct Mfma;
template <> struct Mfma<float> {
  typedef f4 acc;
  static __device__ inline acc zero() { return f4{0.f, 0.f, 0.f, 0.f}; }
  static __device__ inline acc run(float a, float b, acc c) { return mfma4(a, b, c); }
  static __device__ inline int row(int q, int v) { return 4 * q + v; }
};
template <> struct Mfma<double> {
  typedef d4 acc;
  static __device__ inline acc zero() { return d4{0., 0., 0., 0.}; }
  static __device__ inline acc run(double a, double b, acc c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int q, int v) { return 4 * v + q; }
};

// On entry v[k] of lane 16r+i belongs to lane row r; on exit v[r] of lane
// 16k+i holds what was v[k] of lane 16r+i (a 4x4 transpose of lane rows
// against registers).
__device__ inline void row_transpose4(float v[4])
{
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[0]),
                                                  __float_as_uint(v[2]), false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[1]),
                                                  __float_as_uint(v[3]), false, false);
  const auto c = __builtin_amdgcn_permlane16_swap(a[0], b[0], false, false);
  const auto d = __builtin_amdgcn_permlane16_swap(a[1], b[1], false, false);
  v[0] = __uint_as_float(c[0]);
  v[1] = __uint_as_float(c[1]);
  v[2] = __uint_as_float(d[0]);
  v[3] = __uint_as_float(d[1]);
}

// Exact three-piece BF16 split of FP32 values: x = h + m + l, each piece the
// top 8 significant bits of what remains (truncation; the last remainder has
// at most 8 significant bits, so l is exact).  Eight values per operand of
// v_mfma_f32_16x16x32_bf16.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
struct Bf3 { bf16x8 h, m, l; };

// The BF16 pairs (-1, 0) and (0, -1) of the remainder step below, made by
// scalar moves the compiler cannot see through: written as literals, the
// pair (-1, 0) is encoded as the inline constant -1.0 of v_dot2c_f32_bf16,
// which the gfx950 hardware applies as (0, -1).
struct SplitK { bf16x2 n0, n1; };
__device__ inline SplitK split_consts()
{
  unsigned c0, c1;
  asm("s_mov_b32 %0, 0xbf80" : "=s"(c0));
  asm("s_mov_b32 %0, 0xbf800000" : "=s"(c1));
  return SplitK{__builtin_bit_cast(bf16x2, c0), __builtin_bit_cast(bf16x2, c1)};
}

// Each word packs the pieces of two values (element 2p in the low half):
// one byte permute takes the top halves of both, and the remainder of each
// value is one v_dot2c_f32_bf16, x + (-1) * piece, exact because the piece
// is x's own leading bits.  3.5 VALU per value.
__device__ inline Bf3 split3(const float (&v)[8], const SplitK &k)
{
  // in phases over the four words, so that each result is used four
  // instructions after it is written (a dot2c result read by the next
  // instruction costs wait states).  The words are plain scalars until the
  // end: read back from a u32x4 for the dot products, every word came out
  // as word 0 in the compiled code.
  unsigned H[4], M[4], L[4];
  float r[8], t[8];
#pragma unroll
  for (int p = 0; p < 4; ++p)
    H[p] = __builtin_amdgcn_perm(__float_as_uint(v[2 * p + 1]), __float_as_uint(v[2 * p]),
                                 0x07060302u);
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const bf16x2 hp = __builtin_bit_cast(bf16x2, H[p]);
    r[2 * p] = __builtin_amdgcn_fdot2_f32_bf16(hp, k.n0, v[2 * p], false);
    r[2 * p + 1] = __builtin_amdgcn_fdot2_f32_bf16(hp, k.n1, v[2 * p + 1], false);
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
    M[p] = __builtin_amdgcn_perm(__float_as_uint(r[2 * p + 1]), __float_as_uint(r[2 * p]),
                                 0x07060302u);
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const bf16x2 mp = __builtin_bit_cast(bf16x2, M[p]);
    t[2 * p] = __builtin_amdgcn_fdot2_f32_bf16(mp, k.n0, r[2 * p], false);
    t[2 * p + 1] = __builtin_amdgcn_fdot2_f32_bf16(mp, k.n1, r[2 * p + 1], false);
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
    L[p] = __builtin_amdgcn_perm(__float_as_uint(t[2 * p + 1]), __float_as_uint(t[2 * p]),
                                 0x07060302u);
  const u32x4 h = {H[0], H[1], H[2], H[3]}, mm = {M[0], M[1], M[2], M[3]},
              l = {L[0], L[1], L[2], L[3]};
  return Bf3{__builtin_bit_cast(bf16x8, h), __builtin_bit_cast(bf16x8, mm),
             __builtin_bit_cast(bf16x8, l)};
}

// Scheduling pattern: NM matrix instructions, each followed by NV other VALU
// instructions, so that a wavefront issues VALU work (the next group's
// splits) while its own matrix products run.
template <int NM, int NV>
__device__ inline void sched_interleave()
{
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);
  }
}
// D += A.B from the six piece products of order >= 2^-16: with |m| < 2^-7|x|
// and |l| < 2^-14|x| the three dropped ones stay below 2^-20 of each
// product (about 2^-24 in the rms); smallest first
__device__ inline f4 mfma_bf3(const Bf3 &a, const Bf3 &b, f4 d)
{
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.m, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.h, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.m, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, d, 0, 0, 0);
  return d;
}
// Analysis: as xcd_map, but the 8 m of one phase block (m / 8: a phase cache
// line holds the 8 m of a ring, phase_index) go to one XCD, issued next to
// each other for each ring group, so that they share the block's lines in
// that L2; blocks go round-robin over the XCDs.
__device__ inline bool xcd_map_mblock(int id, int nmajor, int nminor, int &major, int &minor)
{
  const int xcd = id & 7, k = id >> 3, per = 8 * nminor;
  const int blk = (k / per) * 8 + xcd, r = k % per;
  minor = r / 8;
  major = blk * 8 + r % 8;
  return major < nmajor;
}

static inline int xcd_grid_mblock(int nmajor, int nminor)
{
  const int nblk = (nmajor + 7) / 8;
  return ((nblk + 7) / 8) * 64 * nminor;
}

constexpr int kFlushBlocks = 8;   // K slices per FP32 partial sum

template <int NCB>
__device__ inline void flush(f4 (&tot)[NCB][4], f4 (&c)[NCB][4])
{
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      tot[cb][r] += c[cb][r];
      c[cb][r] = f4{0.f, 0.f, 0.f, 0.f};
    }
}

// stores N = t1+t2 and S = sg*(t1-t2) of the 64 ring pairs of a wavefront:
// lane n+16q holds rows 16r+4q+v (v = 0..3) of column n of accumulator r
template <int NCB>
__device__ inline void store_phases(const LegArgs &a, int m, int sb, int comp0, int ncol,
                                    double sg, const f4 (&t1)[NCB][4],
                                    const f4 (&t2)[NCB][4])
{
  const int lane = threadIdx.x, n = lane & 15, q = lane >> 4;
  double *ph = reinterpret_cast<double *>(a.phase);
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb) {
    const int col = cb * 16 + n;
    if (col >= ncol) continue;
    const int comp = comp0 + (col >> 1), part = col & 1;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ring = sb * 64 + 16 * r + 4 * q + v;
        if (ring >= a.nslots) continue;
        const double x1 = t1[cb][r][v], x2 = t2[cb][r][v];
        ph[2 * phase_index(comp, m, ring, 0, a.nslots, a.nmb) + part] = x1 + x2;
        if (a.ispair[ring])
          ph[2 * phase_index(comp, m, ring, 1, a.nslots, a.nmb) + part] = sg * (x1 - x2);
      }
  }
}

// ===========================================================================
// spin 0 synthesis
// ===========================================================================
// Memory operations are issued one block ahead of their use: the a_lm rows
// of the next block (vector loads, which complete in order) at the top of a
// block, the recurrence coefficients of the next block (scalar loads, which
// complete out of order and are waited for as a whole) after the first
// two-step of a block, so that the wait for the current coefficients does
// not also wait for the prefetch.
// One wavefront: order m (wave-uniform) for the 64 ring pairs of slot block sb.
// CROW: read the coefficient rows through the constant address space (the
// batch kernels, whose job arguments come from memory; kernel-argument
// pointers get scalar loads without it, and plain pointers schedule better)
template <int NCB, bool CROW = false>
__device__ void msynth_spin0_tile(const LegArgs &a, int m, int sb)
{
  const int lane = threadIdx.x;
  if (m > a.mmax) return;
  const int lmax = a.lmax, nc = a.ncomp_alm, ncol = 2 * a.ntl;
  const int slot = sb * 64 + lane;
  const bool valid = slot < a.nslots;
  const int s = valid ? slot : a.nslots - 1;
  const double cth = a.cth[s], sth = a.sth[s];
  const bool active = valid && (a.mlim[s] >= m);
  const int n = lane & 15, kq = lane >> 4;

  f4 tot1[NCB][4], tot2[NCB][4];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) tot1[cb][r] = tot2[cb][r] = f4{0.f, 0.f, 0.f, 0.f};

  if (wave_any(active)) {
    typedef std::conditional_t<CROW, URow2, const double2 *> Row2;
    const Row2 rf = Row2(reinterpret_cast<const double2 *>(a.tab) + (tab_ofs(m, lmax) - m));
    const int lpad = lmax + kTabPad;   // last readable entry of the row
    // row l of this m, first column of this launch
    const float *arow = reinterpret_cast<const float *>(a.alm) +
                        ((int64_t)(tri_ofs(m, lmax) - m) * nc + a.t0) * 2;
    int colo[NCB];
    bool colok[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int col = cb * 16 + n;
      colok[cb] = col < ncol;
      colo[cb] = colok[cb] ? col : 0;
    }
    // B rows l0+2k (accumulator 1) and l0+2k+1 (accumulator 2).  Rows past
    // lmax (up to lmax+15) are read from the next m, or from the slack rows
    // at the end of the canonical array, and masked where they are used.
    int boff1[NCB], boff2[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      boff1[cb] = 2 * kq * 2 * nc + colo[cb];
      boff2[cb] = boff1[cb] + 2 * nc;
    }
    auto ldB = [&](int l0, float (&b1)[NCB], float (&b2)[NCB]) {
      const float *row = arow + (int64_t)l0 * 2 * nc;
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        b1[cb] = row[boff1[cb]];
        b2[cb] = row[boff2[cb]];
      }
    };

    double lam1 = 0., lam2;
    int scale;
    mypow(sth, m, lam2, scale);
    const double mf = a.mfac[m];
    lam2 *= (m & 1) ? -mf : mf;
    normalize(lam2, scale, kFTol);
    // a ring pair beyond its m limit contributes nothing: zero lambda, in
    // range, so it neither contributes nor holds back the mode switches
    if (!active) { lam2 = 0.; scale = 0; }

    int l = m;
    bool dead = false;
    while (wave_all(scale < 1)) {
      if (l + 2 > lmax) { dead = true; break; }
      const double2 r0 = rf[l], r1 = rf[l + 1];
      lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
      lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
      if (fabs(lam2) > kFTol) { lam1 *= kFSmall; lam2 *= kFSmall; ++scale; }
      l += 2;
    }
    if (!dead) {
      double cf = corfac(scale);
      bool scaled = wave_any(scale < 0);
      if (!scaled) { lam1 *= cf; lam2 *= cf; cf = 1.; }
      f4 c1[NCB][4], c2[NCB][4];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) c1[cb][r] = c2[cb][r] = f4{0.f, 0.f, 0.f, 0.f};
      double2 rc[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) rc[k] = rf[min(l + k, lpad)];
      float bc1[NCB], bc2[NCB];
      ldB(l, bc1, bc2);
      auto mfmas = [&](const float (&v1)[4], const float (&v2)[4], const float (&b1)[NCB],
                       const float (&b2)[NCB]) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            c1[cb][r] = mfma4(v1[r], b1[cb], c1[cb][r]);
            c2[cb][r] = mfma4(v2[r], b2[cb], c2[cb][r]);
          }
      };
      int nblk = 0;
      // one block of 8 l with every check (scaled domain, rows past lmax)
      auto general_block = [&]() {
        const int lb = l;
        float bn1[NCB], bn2[NCB];
        ldB(lb + 8, bn1, bn2);
        double2 rn[8];
        float v1[4], v2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double o0 = 0., o1 = 0.;
          if (l <= lmax) {
            const double2 r0 = rc[2 * q], r1 = rc[2 * q + 1];
            o0 = lam2 * cf;
            lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
            o1 = (l + 1 <= lmax) ? lam1 * cf : 0.;
            lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
            if (scaled && fabs(lam2) > kFTol) {
              lam1 *= kFSmall; lam2 *= kFSmall; ++scale;
              cf = corfac(scale);
            }
          }
          v1[q] = (float)o0;
          v2[q] = (float)o1;
          l += 2;
          if (q == 0) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 8; ++k) rn[k] = rf[min(lb + 8 + k, lpad)];
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        row_transpose4(v1);
        row_transpose4(v2);
        float b1[NCB], b2[NCB];
        {
          const int l1 = lb + 2 * kq, l2 = l1 + 1;
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb) {
            b1[cb] = (colok[cb] && l1 <= lmax) ? (float)bc1[cb] : 0.f;
            b2[cb] = (colok[cb] && l2 <= lmax) ? (float)bc2[cb] : 0.f;
          }
        }
        mfmas(v1, v2, b1, b2);
        if (scaled && wave_all(scale >= 0)) {
          lam1 *= cf; lam2 *= cf; cf = 1.;
          scaled = false;
        }
        if (++nblk == kFlushBlocks) {
          nblk = 0;
          flush(tot1, c1);
          flush(tot2, c2);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) rc[k] = rn[k];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) { bc1[cb] = bn1[cb]; bc2[cb] = bn2[cb]; }
      };

      // 1: scaled domain
      while (scaled && l <= lmax) general_block();
      // 2: unscaled, whole groups of kFlushBlocks blocks, software-pipelined
      // (the matrix products of a block are issued during the recurrence of
      // the next one)
      constexpr int GL = 8 * kFlushBlocks;   // l per group
      if (!scaled && l + GL - 1 <= lmax) {
        float pv1[4] = {0.f, 0.f, 0.f, 0.f}, pv2[4] = {0.f, 0.f, 0.f, 0.f};
        float pb1[NCB], pb2[NCB];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) pb1[cb] = pb2[cb] = 0.f;
        while (l + GL - 1 <= lmax) {
#pragma unroll
          for (int k = 0; k < kFlushBlocks; ++k) {
            const int lb = l;
            float bn1[NCB], bn2[NCB];
            ldB(lb + 8, bn1, bn2);
            double2 rn[8];
            float v1[4], v2[4];
            // first use of this block's coefficients (waits for the scalar
            // loads in flight), then the next block's requests
            const double t00 = cth * lam2, r00x = rc[0].x;
            __builtin_amdgcn_sched_barrier(0);
            // lb + 15 <= lmax + kTabPad in this phase
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) rn[kk] = rf[lb + 8 + kk];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const double2 r0 = rc[2 * q], r1 = rc[2 * q + 1];
              v1[q] = (float)lam2;
              // unit-b table (every b_l = 1): lambda_l / s_l
              lam1 = (q == 0 ? t00 * r00x : (cth * lam2) * r0.x) - lam1;
              v2[q] = (float)lam1;
              lam2 = (cth * lam1) * r1.x - lam2;
              if (q == 1) mfmas(pv1, pv2, pb1, pb2);
            }
            row_transpose4(v1);
            row_transpose4(v2);
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
              pb1[cb] = colok[cb] ? (float)bc1[cb] : 0.f;
              pb2[cb] = colok[cb] ? (float)bc2[cb] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) { pv1[q] = v1[q]; pv2[q] = v2[q]; }
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) rc[kk] = rn[kk];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) { bc1[cb] = bn1[cb]; bc2[cb] = bn2[cb]; }
            l += 8;
          }
          flush(tot1, c1);
          flush(tot2, c2);
        }
        mfmas(pv1, pv2, pb1, pb2);
        nblk = 0;
      }
      // 3: the remaining rows
      while (l <= lmax) general_block();
      flush(tot1, c1);
      flush(tot2, c2);
    }
  }
  store_phases<NCB>(a, m, sb, a.t0, ncol, 1., tot1, tot2);
}

// ===========================================================================
// spin s synthesis (and ALM2MAP_DERIV1)
// ===========================================================================
// CT: 0 FP32 matrix cores; 1 BF16 matrix cores on three-piece splits in the
// unscaled part (groups of 32 l), FP32 for the scaled blocks and the tail
// cst (CT 1): this wavefront's 64 entries of LDS for the coefficient groups
template <int NCB, bool DERIV, int CT, bool CROW = false>
__device__ void msynth_spin_tile(const LegArgs &a, int m, int sb, double2 *cst)
{
  const int lane = threadIdx.x;
  if (m > a.mmax) return;
  const int lmax = a.lmax, nc = a.ncomp_alm, sp = a.spin, ncol = 4 * a.ntl;
  const int slot = sb * 64 + lane;
  const bool valid = slot < a.nslots;
  const int s = valid ? slot : a.nslots - 1;
  const double cth = a.cth[s], sth = a.sth[s];
  const bool active = valid && (a.mlim[s] >= m);
  const SpinStart st = spin_start(m, sp);
  const int n = lane & 15, kq = lane >> 4;

  f4 tot1[NCB][4], tot2[NCB][4];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) tot1[cb][r] = tot2[cb][r] = f4{0.f, 0.f, 0.f, 0.f};

  if (wave_any(active) && st.mhi <= lmax) {
    typedef std::conditional_t<CROW, URow, const double *> Row;
    const Row fxb = Row(a.tab + 3 * (tab_ofs(m, lmax) - m));
    const int lpad = lmax + kTabPad;
    auto fx = [&](int l) {
      const int k = min(l, lpad);
      return Fx{fxb[3 * k], fxb[3 * k + 1], fxb[3 * k + 2]};
    };
    // unclamped: for l <= lpad only
    auto fxu = [&](int l) { return Fx{fxb[3 * l], fxb[3 * l + 1], fxb[3 * l + 2]}; };
    constexpr int NALM = DERIV ? 1 : 2;
    const float *arow = reinterpret_cast<const float *>(a.alm) +
                        ((int64_t)(tri_ofs(m, lmax) - m) * nc + a.t0 * NALM) * 2;
    // Column col = 4t+q is (Q.re, Q.im, U.re, U.im)[q] of transform t.  Its
    // B entry is X = (E.re, E.im, B.re, B.im)[q] for lambda_w and
    // Y = (-B.im, B.re, E.im, -E.re)[q] for lambda_x (libsharp/
    // sharp_core_inc2.c:282-306): offsets into the row and signs (offset 0
    // and sign 0 for unused columns).
    int xo[NCB], yo[NCB];
    float xs[NCB], ys[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int col = cb * 16 + n, t = col >> 2, q = col & 3;
      const bool ok = col < ncol;
      if (!DERIV) {
        xo[cb] = ok ? col : 0;
        xs[cb] = ok ? 1.f : 0.f;
        yo[cb] = ok ? (col ^ 3) : 0;
        ys[cb] = ok ? ((q == 0 || q == 3) ? -1.f : 1.f) : 0.f;
      } else {
        // no curl a_lm: X = (E.re, E.im, 0, 0), Y = (0, 0, E.im, -E.re)
        xo[cb] = ok ? 2 * t + (q & 1) : 0;
        xs[cb] = (ok && q < 2) ? 1.f : 0.f;
        yo[cb] = ok ? 2 * t + (q == 2 ? 1 : 0) : 0;
        ys[cb] = ok ? (q == 2 ? 1.f : (q == 3 ? -1.f : 0.f)) : 0.f;
      }
    }
    // B row l0+k.  Rows past lmax (up to lmax+7) are read from the next m,
    // or from the slack rows at the end of the canonical array, and masked
    // where they are used.
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      xo[cb] += kq * 2 * nc;
      yo[cb] += kq * 2 * nc;
    }
    auto ldB = [&](int l0, float (&bx)[NCB], float (&by)[NCB]) {
      const float *row = arow + (int64_t)l0 * 2 * nc;
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        bx[cb] = row[xo[cb]];
        by[cb] = row[yo[cb]];
      }
    };

    double rec1p = 0., rec1m = 0., rec2p, rec2m;
    int scalep, scalem;
    spin_init(cth, sth, sp, st, a.prefac[m], a.fscale[m], rec2p, rec2m, scalep, scalem);
    // a ring pair beyond its m limit contributes nothing: zero lambda, in
    // range, so it neither contributes nor holds back the mode switches
    if (!active) { rec2p = rec2m = 0.; scalep = scalem = 0; }
    int l = st.mhi;
    bool dead = false;
    while (wave_all(scalep < 1 && scalem < 1)) {
      if (l + 2 > lmax) { dead = true; break; }
      rec_step(rec1p, rec1m, rec2p, rec2m, cth, fx(l + 1));
      rec_step(rec2p, rec2m, rec1p, rec1m, cth, fx(l + 2));
      rescale1(rec1p, rec2p, scalep);
      rescale1(rec1m, rec2m, scalem);
      l += 2;
    }
    if (!dead) {
      double cfp = corfac(scalep), cfm = corfac(scalem);
      bool scaled = wave_any(scalep < 0 || scalem < 0);
      if (!scaled) {
        rec1p *= cfp; rec2p *= cfp; rec1m *= cfm; rec2m *= cfm;
        cfp = cfm = 1.;
      }
      f4 c1[NCB][4], c2[NCB][4];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) c1[cb][r] = c2[cb][r] = f4{0.f, 0.f, 0.f, 0.f};
      Fx fc[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) fc[k] = fx(l + 1 + k);
      float bcx[NCB], bcy[NCB];
      ldB(l, bcx, bcy);
      const bool odd = (kq & 1) != 0;
      auto mfmas = [&](const float (&v1)[4], const float (&v2)[4], const float (&b1)[NCB],
                       const float (&b2)[NCB]) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            c1[cb][r] = mfma4(v1[r], b1[cb], c1[cb][r]);
            c2[cb][r] = mfma4(v2[r], b2[cb], c2[cb][r]);
          }
      };
      // accumulator 1 (N+S side): K rows (w_l, x_l+1, w_l+2, x_l+3) with B
      // rows (X_l, Y_l+1, X_l+2, Y_l+3); accumulator 2 the other way round
      auto bvals = [&](int lb, float (&b1)[NCB], float (&b2)[NCB]) {
        const bool okr = lb + kq <= lmax;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
          const float bx = okr ? xs[cb] * (float)bcx[cb] : 0.f;
          const float by = okr ? ys[cb] * (float)bcy[cb] : 0.f;
          b1[cb] = odd ? by : bx;
          b2[cb] = odd ? bx : by;
        }
      };
      int nblk = 0;
      // one block of 4 l with every check (scaled domain, rows past lmax)
      auto general_block = [&]() {
        const int lb = l;
        float bnx[NCB], bny[NCB];
        ldB(lb + 4, bnx, bny);
        Fx fn[4];
        float v1[4], v2[4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          double lw0 = 0., lx0 = 0., lw1 = 0., lx1 = 0.;
          if (l <= lmax) {
            const double rp0 = rec2p * cfp, rm0 = rec2m * cfm;
            rec_step(rec1p, rec1m, rec2p, rec2m, cth, fc[2 * h]);
            const double rp1 = rec1p * cfp, rm1 = rec1m * cfm;
            rec_step(rec2p, rec2m, rec1p, rec1m, cth, fc[2 * h + 1]);
            if (scaled) {
              const bool a1 = rescale1(rec1p, rec2p, scalep);
              const bool a2 = rescale1(rec1m, rec2m, scalem);
              if (a1 || a2) { cfp = corfac(scalep); cfm = corfac(scalem); }
            }
            lw0 = rp0 + rm0;
            lx0 = rm0 - rp0;
            if (l + 1 <= lmax) {
              lw1 = rp1 + rm1;
              lx1 = rm1 - rp1;
            }
          }
          v1[2 * h] = (float)lw0;
          v1[2 * h + 1] = (float)lx1;
          v2[2 * h] = (float)lx0;
          v2[2 * h + 1] = (float)lw1;
          l += 2;
          if (h == 0) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) fn[k] = fx(lb + 5 + k);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        row_transpose4(v1);
        row_transpose4(v2);
        float b1[NCB], b2[NCB];
        bvals(lb, b1, b2);
        mfmas(v1, v2, b1, b2);
        if (scaled && wave_all(scalep >= 0 && scalem >= 0)) {
          rec1p *= cfp; rec2p *= cfp; rec1m *= cfm; rec2m *= cfm;
          cfp = cfm = 1.;
          scaled = false;
        }
        if (++nblk == kFlushBlocks) {
          nblk = 0;
          flush(tot1, c1);
          flush(tot2, c2);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) fc[k] = fn[k];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) { bcx[cb] = bnx[cb]; bcy[cb] = bny[cb]; }
      };

      // 1: scaled domain
      while (scaled && l <= lmax) general_block();
      if constexpr (CT == 1) {
        // 2 (BF16 splits): whole groups of 32 l.  The K index k = 0..31 of
        // the products is l0+k; accumulator 1 takes (w, x, w, x, ...) with B
        // rows (X, Y, X, Y, ...), accumulator 2 the other way round, as in
        // the FP32 blocks.  Lane n+16kq holds B rows l0+8kq..l0+8kq+7.
        if (!scaled && l + 31 <= lmax) {
          flush(tot1, c1);
          flush(tot2, c2);
          nblk = 0;
          // B from the piece words the gather wrote per output column
          // (bs_pairs): accumulator 1 takes the words of this lane's column
          // c, accumulator 2 those of column c^3, whose products carry the
          // sign s_k of column c (added to the total as s_k times the group
          // sum); unused columns read the zero column.  Rows l0+8kq.. are the
          // 4 words from pair (l0+8kq-mhi)/2 (l - mhi is even here).
          const SplitK sk = split_consts();
          const unsigned *bsm = a.bsplit + a.bs_base[m];
          const int np = bs_pairs(lmax, st.mhi);
          const int czero = a.bs_ncol - 1, cbase = 4 * a.t0;
          const unsigned *px[NCB], *py[NCB];
          float s2[NCB];
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb) {
            const int col = cb * 16 + n, q = col & 3;
            const bool ok = col < ncol;
            px[cb] = bsm + (int64_t)3 * (ok ? cbase + col : czero) * np + 4 * kq;
            py[cb] = bsm + (int64_t)3 * (ok ? cbase + (col ^ 3) : czero) * np + 4 * kq;
            s2[cb] = (q == 0 || q == 3) ? -1.f : 1.f;
          }
          // Recurrence coefficients (f0, g) of l0+1 .. l0+32 in this
          // wavefront's LDS slice, staged one group ahead: lanes 0..31 load
          // the next group's 32 entries with vector loads at the top of a
          // group and store them at its end, so that their latency is
          // covered by a whole group; the group reads them at wave-uniform
          // addresses into VGPRs, the operands of the FMAs (an FP64 FMA
          // takes one scalar operand, so scalar coefficients cost a move).
          const double *fxr = a.tab + 3 * (tab_ofs(m, lmax) - m);
          auto fetch = [&](int l1) {
            const int k = min(l1 + (lane & 31), lpad);
            return make_double2(fxr[3 * k], fxr[3 * k + 1]);
          };
          {
            const double2 c = fetch(l + 1);
            if (lane < 32) cst[lane] = c;
          }
          int cbi = 0;
          while (l + 31 <= lmax) {
            const int l0 = l;
            const int r2 = (l0 - st.mhi) >> 1;
            const double2 cnext = fetch(l0 + 33);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const double2 *cq = cst + 32 * cbi;
            u32x4 PX[NCB][3], PY[NCB][3];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
              for (int q = 0; q < 3; ++q) {
                __builtin_memcpy(&PX[cb][q], px[cb] + q * np + r2, sizeof(u32x4));
                __builtin_memcpy(&PY[cb][q], py[cb] + q * np + r2, sizeof(u32x4));
              }
            float a1[32], a2[32];
#pragma unroll
            for (int k = 0; k < 32; k += 2) {
              const double2 f1 = cq[k], f2 = cq[k + 1];
              const double rp0 = rec2p, rm0 = rec2m;
              rec_step_unit(rec1p, rec1m, rec2p, rec2m, cth, Fx{f1.x, f1.y, 1.});
              const double rp1 = rec1p, rm1 = rec1m;
              rec_step_unit(rec2p, rec2m, rec1p, rec1m, cth, Fx{f2.x, f2.y, 1.});
              // combinations in FP32 of the FP32-rounded chains
              const float p0 = (float)rp0, m0 = (float)rm0, p1 = (float)rp1, m1 = (float)rm1;
              a1[k] = p0 + m0;
              a2[k] = m0 - p0;
              a1[k + 1] = m1 - p1;
              a2[k + 1] = p1 + m1;
              // the next group's coefficients to the other slice (last read
              // in the previous group), a quarter of the recurrence after
              // their load was issued: their registers are free again
              // before the products below
              if (k == 8 && lane < 32) cst[32 * (cbi ^ 1) + lane] = cnext;
            }
            l += 32;
            // ring group r's values at k = 8kq..8kq+7 to lane row kq, slot r
#pragma unroll
            for (int p = 0; p < 8; ++p) {
              float t1[4] = {a1[p], a1[8 + p], a1[16 + p], a1[24 + p]};
              float t2[4] = {a2[p], a2[8 + p], a2[16 + p], a2[24 + p]};
              row_transpose4(t1);
              row_transpose4(t2);
#pragma unroll
              for (int r = 0; r < 4; ++r) { a1[8 * r + p] = t1[r]; a2[8 * r + p] = t2[r]; }
            }
            // accumulator 1: (X row 2d, Y row 2d+1) per word, accumulator 2:
            // (Y row 2d, X row 2d+1) = s_k times the words of column c^3
            Bf3 B1[NCB], B2[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
              B1[cb] = Bf3{__builtin_bit_cast(bf16x8, PX[cb][0]), __builtin_bit_cast(bf16x8, PX[cb][1]),
                           __builtin_bit_cast(bf16x8, PX[cb][2])};
              B2[cb] = Bf3{__builtin_bit_cast(bf16x8, PY[cb][0]), __builtin_bit_cast(bf16x8, PY[cb][1]),
                           __builtin_bit_cast(bf16x8, PY[cb][2])};
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              float v1[8], v2[8];
#pragma unroll
              for (int p = 0; p < 8; ++p) { v1[p] = a1[8 * r + p]; v2[p] = a2[8 * r + p]; }
              const Bf3 A1 = split3(v1, sk), A2 = split3(v2, sk);
              // the group's sum from zero, then one addition to the total
              // (as the FP32 blocks flush every 32 l)
#pragma unroll
              for (int cb = 0; cb < NCB; ++cb) {
                tot1[cb][r] += mfma_bf3(A1, B1[cb], f4{0.f, 0.f, 0.f, 0.f});
                const f4 g2 = mfma_bf3(A2, B2[cb], f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int v = 0; v < 4; ++v) tot2[cb][r][v] = fmaf(s2[cb], g2[v], tot2[cb][r][v]);
              }
            }
            cbi ^= 1;
          }
          // the general blocks of the tail read their coefficients from fc
#pragma unroll
          for (int k = 0; k < 4; ++k) fc[k] = fx(l + 1 + k);
          ldB(l, bcx, bcy);
        }
      } else
      // 2: unscaled, whole groups of kFlushBlocks blocks, software-pipelined:
      // the matrix products of a block are issued during the FP64
      // recurrence of the next one
      if (!scaled && l + 4 * kFlushBlocks - 1 <= lmax) {
        float pv1[4] = {0.f, 0.f, 0.f, 0.f}, pv2[4] = {0.f, 0.f, 0.f, 0.f};
        float pb1[NCB], pb2[NCB];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) pb1[cb] = pb2[cb] = 0.f;
        while (l + 4 * kFlushBlocks - 1 <= lmax) {
#pragma unroll
          for (int k = 0; k < kFlushBlocks; ++k) {
            const int lb = l;
            float bnx[NCB], bny[NCB];
            ldB(lb + 4, bnx, bny);
            Fx fn[4];
            float v1[4], v2[4];
            // The first use of this block's coefficients waits for every
            // scalar load in flight; the next block's are requested right
            // after it, so they have a whole block to arrive.
            const double c0m = fma(cth, fc[0].f0, -fc[0].g), c0p = fma(cth, fc[0].f0, fc[0].g);
            __builtin_amdgcn_sched_barrier(0);
            // lb + 8 <= lmax + kTabPad in this phase
#pragma unroll
            for (int q = 0; q < 4; ++q) fn[q] = fxu(lb + 5 + q);
            __builtin_amdgcn_sched_barrier(0);
            {
              // unit-f2 table: lambda_l / s_l
              const double rp0 = rec2p, rm0 = rec2m;
              rec1p = fma(c0m, rec2p, -rec1p);
              rec1m = fma(c0p, rec2m, -rec1m);
              rec_step_unit(rec2p, rec2m, rec1p, rec1m, cth, fc[1]);
              v1[0] = (float)(rp0 + rm0);
              v2[0] = (float)(rm0 - rp0);
              v2[1] = (float)(rec1p + rec1m);
              v1[1] = (float)(rec1m - rec1p);
            }
            {
              const double rp0 = rec2p, rm0 = rec2m;
              rec_step_unit(rec1p, rec1m, rec2p, rec2m, cth, fc[2]);
              rec_step_unit(rec2p, rec2m, rec1p, rec1m, cth, fc[3]);
              v1[2] = (float)(rp0 + rm0);
              v2[2] = (float)(rm0 - rp0);
              v2[3] = (float)(rec1p + rec1m);
              v1[3] = (float)(rec1m - rec1p);
            }
            mfmas(pv1, pv2, pb1, pb2);
            row_transpose4(v1);
            row_transpose4(v2);
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
              const float bx = xs[cb] * (float)bcx[cb], by = ys[cb] * (float)bcy[cb];
              pb1[cb] = odd ? by : bx;
              pb2[cb] = odd ? bx : by;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) { pv1[q] = v1[q]; pv2[q] = v2[q]; fc[q] = fn[q]; }
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) { bcx[cb] = bnx[cb]; bcy[cb] = bny[cb]; }
            l += 4;
          }
          flush(tot1, c1);
          flush(tot2, c2);
        }
        mfmas(pv1, pv2, pb1, pb2);
        nblk = 0;
      }
      // 3: the remaining rows
      while (l <= lmax) general_block();
      flush(tot1, c1);
      flush(tot2, c2);
    }
  }
  const double sg = ((st.mhi - m + sp) & 1) ? -1. : 1.;
  store_phases<NCB>(a, m, sb, 2 * a.t0, ncol, sg, tot1, tot2);
}

// ===========================================================================
// analysis
// ===========================================================================
constexpr int kLS = 33;    // LDS row stride (floats) of the lambda slices: odd,
                           // so row-wise stores and the transposed MFMA loads
                           // hit distinct banks
constexpr int kTS = 20;    // LDS row stride of the [16 x 16] reduction tiles
// BF16-split contraction: lambda slices as [32 l][64 ring pairs] rows of kRS
// floats.  kRS = 4 mod 64: the compiler merges the 8-byte reads of
// slice_row8 into 16-byte reads, and those of 16 lanes on 16 different rows
// then hit distinct banks (64 banks of 4 bytes); measured 1-2.5% faster than
// the 2 mod 32 stride of the 8-byte reads
constexpr int kRS = 68;
// floats per wavefront slice: the larger of the two layouts
constexpr int kSlice = (32 * kRS > 64 * kLS) ? 32 * kRS : 64 * kLS;
// BF16-split spin-s analysis: lambda_w and lambda_x of a ring pair as one
// 8-byte pair, [32 l][64 ring pairs][w, x] rows of kRS2 floats in the two
// slices of a wavefront (one 8-byte store per l and lane, conflict-free; 16
// lanes on 16 rows read 16 bytes each on distinct banks at 4 mod 64)
constexpr int kRS2 = 132;
static_assert(32 * kRS2 <= 2 * kSlice, "paired lambda rows do not fit two slices");

// eight lambda values of row ll, ring pairs c0..c0+7, of a [l][ring] slice
__device__ inline void slice_row8(const float *sl, int ll, int c0, float (&v)[8])
{
  const float2 *p = reinterpret_cast<const float2 *>(sl + ll * kRS + c0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float2 t = p[q];
    v[2 * q] = t.x;
    v[2 * q + 1] = t.y;
  }
}

// eight (lambda_w, lambda_x) pairs of row ll, ring pairs c0..c0+7, of a
// paired [l][ring][w, x] slice
__device__ inline void slice_row8_wx(const float *sl, int ll, int c0, float (&w)[8], float (&x)[8])
{
  const float4 *p = reinterpret_cast<const float4 *>(sl + ll * kRS2 + 2 * c0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 t = p[q];
    w[2 * q] = t.x;
    x[2 * q] = t.y;
    w[2 * q + 1] = t.z;
    x[2 * q + 1] = t.w;
  }
}

// Adds the wavefronts' tiles (in slice 0 of each wavefront) in a fixed order
// and stores rows lc+x+2i, columns cb*16+n of the group's partial a_lm; a
// thread takes 4 adjacent columns of one row (16-byte LDS reads, one vector
// store where all 4 are columns of the launch).
template <typename T, int NT>
__device__ inline void reduce_store(T *const *slices, int nw, T *prow, int64_t ld, int lc,
                                    int lmax, int ncol)
{
  for (int e = threadIdx.x; e < NT * 64; e += blockDim.x) {
    const int tile = e >> 6, i = (e >> 2) & 15, nn = 4 * (e & 3);
    const int off = (tile * 16 + i) * kTS + nn;
    T sum[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) sum[k] = slices[0][off + k];
    for (int ww = 1; ww < nw; ++ww)
#pragma unroll
      for (int k = 0; k < 4; ++k) sum[k] += slices[ww][off + k];
    const int x = tile & 1, cb = tile >> 1;
    const int ll = lc + x + 2 * i, col = cb * 16 + nn;
    if (ll > lmax || col >= ncol) continue;
    T *dst = prow + (int64_t)ll * ld + col;
    if (col + 3 < ncol) {
#pragma unroll
      for (int k = 0; k < 4; ++k) dst[k] = sum[k];
    } else {
      for (int k = 0; k < ncol - col; ++k) dst[k] = sum[k];
    }
  }
}

template <typename T, int NCB>
__device__ inline void write_tiles(T *slice, const typename Mfma<T>::acc (&t)[2][NCB])
{
  const int lane = threadIdx.x & 63, n = lane & 15, q = lane >> 4;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        slice[((cb * 2 + x) * 16 + Mfma<T>::row(q, v)) * kTS + n] = t[x][cb][v];
}

// CT: 0 the matrix cores of T (FP32 or FP64); 1 (T = float) the BF16
// matrix cores on exact three-piece splits of FP32 operands
// W wavefronts per workgroup: ring group g is slots g*64W .. g*64W+64W-1.
template <typename T, int NCB, int CT, int W>
__device__ void manal_spin0_tile(const LegArgs &a, int m, int g,
                                 T (&lds)[W][kSlice], double (&cbuf)[2][64])
{
  static_assert(CT == 0 || sizeof(T) == 4, "BF16 splits are of FP32 values");
  typedef Mfma<T> MM;
  static_assert(2 * NCB * 16 * kTS <= 64 * kLS, "tiles do not fit the slice");
  if (a.gmax[g] < m) return;   // no ring pair of this group reaches m
  const SplitK sk = split_consts();
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lmax = a.lmax, ncol = 2 * a.ntl, nslots = a.nslots;
  const int wbase = g * 64 * W + w * 64;
  const int slot = wbase + lane;
  const bool valid = slot < nslots;
  const int s = valid ? slot : nslots - 1;
  const bool active = valid && (a.mlim[s] >= m);
  const int n = lane & 15, kq = lane >> 4;
  const double *ph = reinterpret_cast<const double *>(a.phase);
  T *prow = static_cast<T *>(a.part) + (int64_t)g * a.part_rows * a.part_ld +
            (int64_t)(tri_ofs(m, lmax) - m - a.part_row0) * a.part_ld + a.part_col0;

  // B: p1 = N+S and p2 = N-S of a ring pair rs, column n
  auto bpair = [&](int rs, int cb, double &p1, double &p2) {
    const bool ok = rs < nslots && a.mlim[rs < nslots ? rs : 0] >= m;
    const bool pr = ok && a.ispair[rs] != 0;
    const int col = cb * 16 + n;
    double nv = 0., sv = 0.;
    if (ok && col < ncol) {
      const int comp = a.t0 + (col >> 1), part = col & 1;
      nv = ph[2 * phase_index(comp, m, rs, 0, nslots, a.nmb) + part];
      if (pr) sv = ph[2 * phase_index(comp, m, rs, 1, nslots, a.nmb) + part];
    }
    p1 = nv + sv;
    p2 = nv - sv;
  };
  // CT 0: ring pairs wbase+4j+kq (K = 4 per product);
  // CT 1: ring pairs wbase+32kb+8kq+jj, split (K = 32 per product)
  T bp[2][CT ? 1 : 16][NCB];
  Bf3 bq[2][CT ? 2 : 1][NCB];
  if constexpr (CT == 0) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        double p1, p2;
        bpair(wbase + 4 * j + kq, cb, p1, p2);
        bp[0][j][cb] = (T)p1;
        bp[1][j][cb] = (T)p2;
      }
  } else {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        float v1[8], v2[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          double p1, p2;
          bpair(wbase + 32 * kb + 8 * kq + jj, cb, p1, p2);
          v1[jj] = (float)p1;
          v2[jj] = (float)p2;
        }
        bq[0][kb][cb] = split3(v1, sk);
        bq[1][kb][cb] = split3(v2, sk);
      }
  }
  // local rows 2j, 2j+1 (0..31) of this lane's ring pair: as one 8-byte
  // pair in row j of a [16][64 ring pairs][2] slice (BF16 splits; rows of
  // kRS2 floats), else at lane * kLS + row
  auto put = [&](int j, T v0, T v1) {
    if constexpr (CT == 1) {
      *reinterpret_cast<float2 *>(lds[w] + j * kRS2 + 2 * lane) = make_float2(v0, v1);
    } else {
      lds[w][lane * kLS + 2 * j] = v0;
      lds[w][lane * kLS + 2 * j + 1] = v1;
    }
  };
  static_assert(16 * kRS2 <= kSlice, "paired lambda rows do not fit the slice");

  const double cth = a.cth[s], sth = a.sth[s];
  const double2 *rf = reinterpret_cast<const double2 *>(a.tab) + (tab_ofs(m, lmax) - m);
  const int lpad = lmax + kTabPad;
  double lam1 = 0., lam2;
  int scale;
  mypow(sth, m, lam2, scale);
  const double mf = a.mfac[m];
  lam2 *= (m & 1) ? -mf : mf;
  normalize(lam2, scale, kFTol);
  // a ring pair beyond its m limit contributes nothing: zero lambda, in
  // range, so it neither contributes nor holds back the mode switches
  if (!active) { lam2 = 0.; scale = 0; }
  // 0: advancing below the IEEE range, 1: scaled, 2: unscaled, 3: idle
  int mode = !wave_any(active) ? 3 : (wave_all(scale < 1) ? 0 : 1);
  mode = __builtin_amdgcn_readfirstlane(mode);
  double cf = corfac(scale);
  int l = m;
  T *mine = lds[w];
  T *slices[W];
#pragma unroll
  for (int ww = 0; ww < W; ++ww) slices[ww] = lds[ww];
  // coefficients rf[lc .. lc+31] of a chunk (64 doubles) in LDS, shared by
  // the workgroup's wavefronts (one m) and staged one chunk ahead; the
  // unscaled chunks read them with wave-uniform addresses
  const double *rfd = reinterpret_cast<const double *>(rf);
  const int rdmax = 2 * lpad + 1;
  auto stage_coef = [&](int lc, double *dst) {
    if (threadIdx.x < 64) dst[threadIdx.x] = rfd[min(2 * lc + (int)threadIdx.x, rdmax)];
  };
  stage_coef(m, cbuf[0]);
  __syncthreads();
  double2 rc[8];

  const int nchunks = (lmax - m + 32) / 32;
  for (int c = 0; c < nchunks; ++c) {
    const int lc = m + 32 * c;
    // the other buffer was last read in chunk c-1, before its barriers
    stage_coef(lc + 32, cbuf[(c + 1) & 1]);
    const double *cq = cbuf[c & 1];
    bool any = false;
    if (mode == 2 && lc + 31 <= lmax) {
      // unscaled and whole: no checks; lambda of every lane stays in range
      any = true;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const double o0 = lam2;
        if constexpr (sizeof(T) == 4) {
          // unit-b table (every b_l = 1): lambda_l / s_l
          lam1 = (cth * lam2) * cq[4 * q] - lam1;
        } else {
          lam1 = (cth * lam2) * cq[4 * q] - lam1 * cq[4 * q + 1];
        }
        const double o1 = lam1;
        if constexpr (sizeof(T) == 4) {
          lam2 = (cth * lam1) * cq[4 * q + 2] - lam2;
        } else {
          lam2 = (cth * lam1) * cq[4 * q + 2] - lam2 * cq[4 * q + 3];
        }
        put(q, (T)o0, (T)o1);
      }
      l += 32;
    } else {
    // recurrence coefficients of the current 8 l; the next 8 are fetched
    // after the first two-step of each group of 8 (see the synthesis kernels)
#pragma unroll
    for (int k = 0; k < 8; ++k) rc[k] = rf[min(l + k, lpad)];
#pragma unroll
    for (int g8 = 0; g8 < 4; ++g8) {
      const int lb = l;
      double2 rn[8];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = 4 * g8 + q;
        double o0 = 0., o1 = 0.;
        if (mode != 3 && l <= lmax) {
          const double2 r0 = rc[2 * q], r1 = rc[2 * q + 1];
          if (mode == 0) {
            lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
            lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
            if (fabs(lam2) > kFTol) { lam1 *= kFSmall; lam2 *= kFSmall; ++scale; }
            if (wave_any(scale >= 1)) { mode = 1; cf = corfac(scale); }
          } else {
            any = true;
            o0 = lam2 * cf;
            lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
            o1 = (l + 1 <= lmax) ? lam1 * cf : 0.;
            lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
            if (mode == 1) {
              if (fabs(lam2) > kFTol) {
                lam1 *= kFSmall; lam2 *= kFSmall; ++scale;
                cf = corfac(scale);
              }
              if (wave_all(scale >= 0)) {
                lam1 *= cf; lam2 *= cf; cf = 1.;
                mode = 2;
              }
            }
          }
        }
        put(j, active ? (T)o0 : (T)0, active ? (T)o1 : (T)0);
        l += 2;
        if (q == 0) {
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int k = 0; k < 8; ++k) rn[k] = rf[min(lb + 8 + k, lpad)];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) rc[k] = rn[k];
    }
    mode = __builtin_amdgcn_readfirstlane(mode);
    }
    // the slice is read back by the same wavefront: LDS operations of one
    // wavefront complete in order, the fence keeps the compiler from moving
    // the loads above the stores
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    typename MM::acc t[2][NCB];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) t[x][cb] = MM::zero();
    if (any) {
      if constexpr (CT == 0) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            const T A = mine[(4 * j + kq) * kLS + x + 2 * n];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) t[x][cb] = MM::run(A, bp[x][j][cb], t[x][cb]);
          }
      } else {
        // A: rows x+2n (l = lc+x+2n), ring pairs 32kb+8kq..+7, for the
        // groups q = (kb, x) in turn (one read gives both x of a kb); the
        // LDS reads and the split of group q+1 are interleaved with the
        // matrix products of group q
        const float *sm = reinterpret_cast<const float *>(mine);
        float a0[8], a1[8];
        slice_row8_wx(sm, n, 8 * kq, a0, a1);
        Bf3 A = split3(a0, sk);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int kb = q >> 1, x = q & 1;
          Bf3 nA = A;
          if (q == 1) slice_row8_wx(sm, n, 32 + 8 * kq, a0, a1);
          if (q < 3) nA = split3((q & 1) ? a0 : a1, sk);
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb) t[x][cb] = mfma_bf3(A, bq[x][kb][cb], t[x][cb]);
          if (q < 3) sched_interleave<6 * NCB, 28 / (6 * NCB) + 1>();
          A = nA;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    write_tiles<T, NCB>(mine, t);
    __syncthreads();
    reduce_store<T, 2 * NCB>(slices, W, prow, a.part_ld, lc, lmax, ncol);
    __syncthreads();
  }
}

template <typename T, int NCB, int CT, int W>
__device__ void manal_spin_tile(const LegArgs &a, int m, int g,
                                T (&lds)[W][2][kSlice], double (&cbuf)[2][96])
{
  static_assert(CT == 0 || sizeof(T) == 4, "BF16 splits are of FP32 values");
  typedef Mfma<T> MM;
  static_assert(2 * NCB * 16 * kTS <= 64 * kLS, "tiles do not fit the slice");
  if (a.gmax[g] < m) return;
  const SplitK sk = split_consts();
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lmax = a.lmax, ncol = 4 * a.ntl, nslots = a.nslots, sp = a.spin;
  const SpinStart st = spin_start(m, sp);
  const int wbase = g * 64 * W + w * 64;
  const int slot = wbase + lane;
  const bool valid = slot < nslots;
  const int s = valid ? slot : nslots - 1;
  const bool active = valid && (a.mlim[s] >= m);
  const int n = lane & 15, kq = lane >> 4;
  const double *ph = reinterpret_cast<const double *>(a.phase);
  T *prow = static_cast<T *>(a.part) + (int64_t)g * a.part_rows * a.part_ld +
            (int64_t)(tri_ofs(m, lmax) - m - a.part_row0) * a.part_ld + a.part_col0;

  // rows m <= l < mhi have no lambda: zero them
  {
    const int nz = ((st.mhi < lmax + 1 ? st.mhi : lmax + 1) - m) * ncol;
    for (int e = threadIdx.x; e < nz; e += blockDim.x)
      prow[(int64_t)(m + e / ncol) * a.part_ld + e % ncol] = (T)0;
  }
  if (st.mhi > lmax) return;
  const double sg = ((st.mhi - m + sp) & 1) ? -1. : 1.;

  // B: P[0] = N + sg*S and P[1] = N - sg*S (columns Q.re, Q.im, U.re, U.im)
  // of a ring pair rs
  auto bpair = [&](int rs, int cb, double &p1, double &p2) {
    const bool ok = rs < nslots && a.mlim[rs < nslots ? rs : 0] >= m;
    const bool pr = ok && a.ispair[rs] != 0;
    const int col = cb * 16 + n;
    double nv = 0., sv = 0.;
    if (ok && col < ncol) {
      const int comp = 2 * a.t0 + (col >> 1), part = col & 1;
      nv = ph[2 * phase_index(comp, m, rs, 0, nslots, a.nmb) + part];
      if (pr) sv = sg * ph[2 * phase_index(comp, m, rs, 1, nslots, a.nmb) + part];
    }
    p1 = nv + sv;
    p2 = nv - sv;
  };
  // CT 0: ring pairs wbase+4j+kq; CT 1: ring pairs wbase+32kb+8kq+jj, split
  T bp[2][CT ? 1 : 16][NCB];
  Bf3 bq[2][CT ? 2 : 1][NCB];
  if constexpr (CT == 0) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        double p1, p2;
        bpair(wbase + 4 * j + kq, cb, p1, p2);
        bp[0][j][cb] = (T)p1;
        bp[1][j][cb] = (T)p2;
      }
  } else {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        float v1[8], v2[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          double p1, p2;
          bpair(wbase + 32 * kb + 8 * kq + jj, cb, p1, p2);
          v1[jj] = (float)p1;
          v2[jj] = (float)p2;
        }
        bq[0][kb][cb] = split3(v1, sk);
        bq[1][kb][cb] = split3(v2, sk);
      }
  }
  auto sidx = [&](int ll) { return CT ? ll * kRS + lane : lane * kLS + ll; };
  // column 4t+q of the lambda_x product takes P's column 4t+3-q with this sign
  // (a(E,B) += lw*P[x] + lx*rot(P[x^1]), libsharp/sharp_core_inc2.c:339-364)
  const T rsgn = ((n & 3) == 0 || (n & 3) == 3) ? (T)-1 : (T)1;

  const double cth = a.cth[s], sth = a.sth[s];
  const double *fxb = a.tab + 3 * (tab_ofs(m, lmax) - m);
  const int lpad = lmax + kTabPad;
  auto fx = [&](int l) {
    const int k = min(l, lpad);
    return Fx{fxb[3 * k], fxb[3 * k + 1], fxb[3 * k + 2]};
  };
  double rec1p = 0., rec1m = 0., rec2p, rec2m;
  int scalep, scalem;
  spin_init(cth, sth, sp, st, a.prefac[m], a.fscale[m], rec2p, rec2m, scalep, scalem);
  // a ring pair beyond its m limit: zero lambda, in range (as for spin 0)
  if (!active) { rec2p = rec2m = 0.; scalep = scalem = 0; }
  int mode = !wave_any(active) ? 3 : (wave_all(scalep < 1 && scalem < 1) ? 0 : 1);
  mode = __builtin_amdgcn_readfirstlane(mode);
  double cfp = corfac(scalep), cfm = corfac(scalem);
  int l = st.mhi;
  T *mw = lds[w][0], *mx = lds[w][1];
  // lambda_w, lambda_x of row ll of this lane's ring pair
  auto put = [&](int ll, T wv, T xv) {
    if constexpr (CT == 1) {
      *reinterpret_cast<float2 *>(mw + ll * kRS2 + 2 * lane) = make_float2(wv, xv);
    } else {
      mw[sidx(ll)] = wv;
      mx[sidx(ll)] = xv;
    }
  };
  T *slices[W];
#pragma unroll
  for (int ww = 0; ww < W; ++ww) slices[ww] = lds[ww][0];

  // Whole chunks below the IEEE range for every wavefront of the workgroup
  // (mode 0, or idle): their lambda are all zero, so their recurrence runs
  // here without stores, products or reduction, and their partial rows are
  // written as zeros.  The workgroup votes after each chunk (two barriers);
  // a chunk that any wavefront leaves in range is rolled back and processed
  // by the chunk loop.  The recurrence steps are those of the chunk loop's
  // mode 0, so the state it continues from is the same.
  const int nchunks = (lmax - st.mhi + 32) / 32;
  int c0 = 0;
  {
    __shared__ int vote[2];
    if (threadIdx.x == 0) vote[0] = vote[1] = 1;
    __syncthreads();
    for (; st.mhi + 32 * c0 + 31 <= lmax; ++c0) {
      const int lc = st.mhi + 32 * c0;
      const double s1p = rec1p, s1m = rec1m, s2p = rec2p, s2m = rec2m;
      const int ssp = scalep, ssm = scalem;
      bool stay = mode == 3;
      if (mode == 0) {
        for (int k = 0; k < 32; k += 2) {
          rec_step(rec1p, rec1m, rec2p, rec2m, cth, fx(lc + 1 + k));
          rec_step(rec2p, rec2m, rec1p, rec1m, cth, fx(lc + 2 + k));
          rescale1(rec1p, rec2p, scalep);
          rescale1(rec1m, rec2m, scalem);
        }
        stay = wave_all(scalep < 1 && scalem < 1);
      }
      if (!stay && lane == 0) vote[c0 & 1] = 0;
      __syncthreads();
      const bool all = vote[c0 & 1] != 0;
      if (threadIdx.x == 0) vote[(c0 + 1) & 1] = 1;
      __syncthreads();
      if (!all) {
        rec1p = s1p; rec1m = s1m; rec2p = s2p; rec2m = s2m;
        scalep = ssp; scalem = ssm;
        break;
      }
    }
    const int nz = 32 * c0 * ncol;
    for (int e = threadIdx.x; e < nz; e += blockDim.x)
      prow[(int64_t)(st.mhi + e / ncol) * a.part_ld + e % ncol] = (T)0;
  }
  l = st.mhi + 32 * c0;

  // coefficients fx(lc+1 .. lc+32) of a chunk (96 doubles) in LDS, shared
  // by the workgroup's wavefronts (one m) and staged one chunk ahead; the
  // unscaled chunks read them with wave-uniform addresses
  const int fdmax = 3 * lpad + 2;
  auto stage_coef = [&](int lc, double *dst) {
    if (threadIdx.x < 96) dst[threadIdx.x] = fxb[min(3 * (lc + 1) + (int)threadIdx.x, fdmax)];
  };
  stage_coef(l, cbuf[c0 & 1]);
  __syncthreads();
  Fx fc[4];

  for (int c = c0; c < nchunks; ++c) {
    const int lc = st.mhi + 32 * c;
    // the next chunk's coefficients: loaded here, stored into the other
    // buffer (last read in chunk c-1, before its barriers) at the end of the
    // chunk, so that the load's latency is not waited for up front
    const double cnext =
        threadIdx.x < 96 ? fxb[min(3 * (lc + 33) + (int)threadIdx.x, fdmax)] : 0.;
    // the buffer index as a VGPR value (an identity DPP move, which the
    // compiler does not take as uniform): the reads below then address the
    // buffer as one VGPR plus immediate offsets, not as a scalar address
    // moved into a VGPR for every read
    const double *cq = &cbuf[0][0] + 96 * __builtin_amdgcn_mov_dpp(c & 1, 0xE4, 0xF, 0xF, false);
    bool any = false;
    if (mode == 2 && lc + 31 <= lmax) {
      // unscaled and whole: no checks; lambda of every lane stays in range
      any = true;
      // the coefficients of two-step j+1 are read from LDS during two-step
      // j (the unit-f2 forms need f0 and g only)
      constexpr bool kUnit = sizeof(T) == 4;
      auto ldq = [&](int j, Fx &f1, Fx &f2) {
        f1.f0 = cq[6 * j];
        f1.g = cq[6 * j + 1];
        f1.f2 = kUnit ? 1. : cq[6 * j + 2];
        f2.f0 = cq[6 * j + 3];
        f2.g = cq[6 * j + 4];
        f2.f2 = kUnit ? 1. : cq[6 * j + 5];
      };
      Fx q1, q2;
      ldq(0, q1, q2);
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const Fx f1 = q1, f2 = q2;
        if (j < 15) ldq(j + 1, q1, q2);
        const double rp0 = rec2p, rm0 = rec2m;
        double rp1, rm1;
        if constexpr (sizeof(T) == 4) {
          // unit-f2 table (the mixed forms): lambda_l / s_l
          rec_step_unit(rec1p, rec1m, rec2p, rec2m, cth, f1);
          rp1 = rec1p; rm1 = rec1m;
          rec_step_unit(rec2p, rec2m, rec1p, rec1m, cth, f2);
        } else {
          rec_step(rec1p, rec1m, rec2p, rec2m, cth, f1);
          rp1 = rec1p; rm1 = rec1m;
          rec_step(rec2p, rec2m, rec1p, rec1m, cth, f2);
        }
        if constexpr (sizeof(T) == 4) {
          // combinations in FP32 of the FP32-rounded chains
          const float p0 = (float)rp0, m0 = (float)rm0, p1 = (float)rp1, m1 = (float)rm1;
          put(2 * j, p0 + m0, m0 - p0);
          put(2 * j + 1, p1 + m1, m1 - p1);
        } else {
          mw[sidx(2 * j)] = (T)(rp0 + rm0);
          mw[sidx(2 * j + 1)] = (T)(rp1 + rm1);
          mx[sidx(2 * j)] = (T)(rm0 - rp0);
          mx[sidx(2 * j + 1)] = (T)(rm1 - rp1);
        }
      }
      l += 32;
    } else {
    // coefficients fx(l+1..l+4) of the current two two-steps, prefetched as
    // in the synthesis kernels
#pragma unroll
    for (int k = 0; k < 4; ++k) fc[k] = fx(l + 1 + k);
#pragma unroll
    for (int g4 = 0; g4 < 8; ++g4) {
      const int lb = l;
      Fx fn[4];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = 2 * g4 + h;
        double lw0 = 0., lx0 = 0., lw1 = 0., lx1 = 0.;
        if (mode != 3 && l <= lmax) {
          if (mode == 0) {
            rec_step(rec1p, rec1m, rec2p, rec2m, cth, fc[2 * h]);
            rec_step(rec2p, rec2m, rec1p, rec1m, cth, fc[2 * h + 1]);
            rescale1(rec1p, rec2p, scalep);
            rescale1(rec1m, rec2m, scalem);
            if (wave_any(scalep >= 1 || scalem >= 1)) {
              mode = 1;
              cfp = corfac(scalep); cfm = corfac(scalem);
            }
          } else {
            any = true;
            const double rp0 = rec2p * cfp, rm0 = rec2m * cfm;
            rec_step(rec1p, rec1m, rec2p, rec2m, cth, fc[2 * h]);
            const double rp1 = rec1p * cfp, rm1 = rec1m * cfm;
            rec_step(rec2p, rec2m, rec1p, rec1m, cth, fc[2 * h + 1]);
            lw0 = rp0 + rm0;
            lx0 = rm0 - rp0;
            if (l + 1 <= lmax) {
              lw1 = rp1 + rm1;
              lx1 = rm1 - rp1;
            }
            if (mode == 1) {
              const bool a1 = rescale1(rec1p, rec2p, scalep);
              const bool a2 = rescale1(rec1m, rec2m, scalem);
              if (a1 || a2) { cfp = corfac(scalep); cfm = corfac(scalem); }
              if (wave_all(scalep >= 0 && scalem >= 0)) {
                rec1p *= cfp; rec2p *= cfp; rec1m *= cfm; rec2m *= cfm;
                cfp = cfm = 1.;
                mode = 2;
              }
            }
          }
        }
        put(2 * j, active ? (T)lw0 : (T)0, active ? (T)lx0 : (T)0);
        put(2 * j + 1, active ? (T)lw1 : (T)0, active ? (T)lx1 : (T)0);
        l += 2;
        if (h == 0) {
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int k = 0; k < 4; ++k) fn[k] = fx(lb + 5 + k);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) fc[k] = fn[k];
    }
    mode = __builtin_amdgcn_readfirstlane(mode);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    typename MM::acc tw[2][NCB], tx[2][NCB];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) tw[x][cb] = tx[x][cb] = MM::zero();
    if (any) {
      if constexpr (CT == 0) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            const int off = (4 * j + kq) * kLS + x + 2 * n;
            const T Aw = mw[off], Ax = mx[off];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
              tw[x][cb] = MM::run(Aw, bp[x][j][cb], tw[x][cb]);
              tx[x][cb] = MM::run(Ax, bp[x ^ 1][j][cb], tx[x][cb]);
            }
          }
      } else {
        // groups q = (x, kb) in turn; the LDS reads and the splits of group
        // q+1 are interleaved with the matrix products of group q
        const float *sw = reinterpret_cast<const float *>(mw);
        Bf3 Aw, Ax;
        {
          float aw[8], ax[8];
          slice_row8_wx(sw, 2 * n, 8 * kq, aw, ax);
          Aw = split3(aw, sk);
          Ax = split3(ax, sk);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int x = q >> 1, kb = q & 1;
          Bf3 nw = Aw, nx = Ax;
          if (q < 3) {
            const int x1 = (q + 1) >> 1, kb1 = (q + 1) & 1;
            float aw[8], ax[8];
            slice_row8_wx(sw, x1 + 2 * n, 32 * kb1 + 8 * kq, aw, ax);
            nw = split3(aw, sk);
            nx = split3(ax, sk);
          }
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb) {
            tw[x][cb] = mfma_bf3(Aw, bq[x][kb][cb], tw[x][cb]);
            tx[x][cb] = mfma_bf3(Ax, bq[x ^ 1][kb][cb], tx[x][cb]);
          }
          if (q < 3) sched_interleave<12 * NCB, 56 / (12 * NCB) + 1>();
          Aw = nw;
          Ax = nx;
        }
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if constexpr (sizeof(T) == 4) {
              // lane i ^ 3 of the quad by a DPP quad permute [3,2,1,0]
              const float t = tx[x][cb][v];
              const float y = __int_as_float(
                  __builtin_amdgcn_mov_dpp(__float_as_int(t), 0x1B, 0xF, 0xF, false));
              tw[x][cb][v] = fmaf(rsgn, y, tw[x][cb][v]);
            } else {
              tw[x][cb][v] += rsgn * __shfl_xor(tx[x][cb][v], 3);
            }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    write_tiles<T, NCB>(mw, tw);
    if (threadIdx.x < 96) cbuf[(c + 1) & 1][threadIdx.x] = cnext;
    __syncthreads();
    reduce_store<T, 2 * NCB>(slices, W, prow, a.part_ld, lc, lmax, ncol);
    __syncthreads();
  }
}

// ===========================================================================
// kernels: one job (grid over its tiles), or a batch (persistent grid over a
// work list of several jobs)
// ===========================================================================
// wave-uniform index of this wavefront in its workgroup (the synthesis
// workgroups are blockDim.y wavefronts of 64 lanes, one m each)
__device__ inline int wave_y() { return __builtin_amdgcn_readfirstlane(threadIdx.y); }

template <int NCB>
__global__ void __launch_bounds__(64 * MB) k_msynth_spin0(LegArgs a, int nsb)
{
  int mb, sb;
  if (xcd_map(blockIdx.x, a.nmb, nsb, mb, sb)) msynth_spin0_tile<NCB>(a, mb * MB + wave_y(), sb);
}

template <int NCB, bool DERIV, int CT>
__global__ void __launch_bounds__(64 * MB) k_msynth_spin(LegArgs a, int nsb)
{
  __shared__ double2 cst[CT ? MB : 1][64];
  int mb, sb;
  if (xcd_map(blockIdx.x, a.nmb, nsb, mb, sb))
    msynth_spin_tile<NCB, DERIV, CT>(a, mb * MB + wave_y(), sb, cst[CT ? wave_y() : 0]);
}

// minimum waves per SIMD: the FP32/BF16 spin-0 analysis with one column block
// uses 36 KB of LDS per workgroup, so four workgroups fit a CU when it keeps
// to 128 VGPRs (the BF16 form needs 134 unbounded: 3 per CU, 11% slower)
template <typename T, int NCB>
constexpr int kAnal0MinWaves = (sizeof(T) == 4 && NCB == 1) ? 4 : 1;

template <typename T, int NCB, int CT>
__global__ void __launch_bounds__(64 * kMixWaves0, (kAnal0MinWaves<T, NCB>)) k_manal_spin0(LegArgs a)
{
  __shared__ __align__(16) T lds[kMixWaves0][kSlice];
  __shared__ double cbuf[2][64];
  int m, g;
  if (xcd_map_mblock(blockIdx.x, a.m1 - a.m0, a.ngroups, m, g))
    manal_spin0_tile<T, NCB, CT, kMixWaves0>(a, a.m0 + m, g, lds, cbuf);
}

template <typename T, int NCB, int CT>
__global__ void __launch_bounds__(64 * kMixWavesS) k_manal_spin(LegArgs a)
{
  __shared__ __align__(16) T lds[kMixWavesS][2][kSlice];
  __shared__ double cbuf[2][96];
  int m, g;
  if (xcd_map_mblock(blockIdx.x, a.m1 - a.m0, a.ngroups, m, g))
    manal_spin_tile<T, NCB, CT, kMixWavesS>(a, a.m0 + m, g, lds, cbuf);
}

// Batch forms (for_flat_item / for_group_items): synthesis items are
// (m-block of blockDim.y orders m, slot block) tiles, one order m per
// wavefront (a workgroup then writes whole phase lines); analysis items are
// (m, ring group of W wavefronts) tiles.
template <int NCB, bool FLAT>
__global__ void __launch_bounds__(64 * MB) kb_msynth_spin0(const LegArgs *__restrict__ jobs,
                                                           const BatchItem *__restrict__ items,
                                                           const int *__restrict__ xofs,
                                                           int *__restrict__ ctr)
{
  auto tile = [&](const BatchItem &it) {
    msynth_spin0_tile<NCB, true>(job_args(jobs, it), it.major * blockDim.y + wave_y(), it.minor);
  };
  if constexpr (FLAT) {
    for_flat_item(items, xofs, tile);
  } else {
    __shared__ int fetched;
    for_group_items(items, xofs, ctr, fetched, tile);
  }
}

template <int NCB, bool DERIV, int CT, bool FLAT>
__global__ void __launch_bounds__(64 * MB) kb_msynth_spin(const LegArgs *__restrict__ jobs,
                                                          const BatchItem *__restrict__ items,
                                                          const int *__restrict__ xofs,
                                                          int *__restrict__ ctr)
{
  __shared__ double2 cst[CT ? MB : 1][64];
  auto tile = [&](const BatchItem &it) {
    msynth_spin_tile<NCB, DERIV, CT, true>(job_args(jobs, it), it.major * blockDim.y + wave_y(),
                                           it.minor, cst[CT ? wave_y() : 0]);
  };
  if constexpr (FLAT) {
    for_flat_item(items, xofs, tile);
  } else {
    __shared__ int fetched;
    for_group_items(items, xofs, ctr, fetched, tile);
  }
}

template <typename T, int NCB, int CT, int W>
__global__ void __launch_bounds__(64 * W, (kAnal0MinWaves<T, NCB>)) kb_manal_spin0(const LegArgs *__restrict__ jobs,
                                                         const BatchItem *__restrict__ items,
                                                         const int *__restrict__ xofs, int *__restrict__ ctr)
{
  __shared__ __align__(16) T lds[W][kSlice];
  __shared__ double cbuf[2][64];
  __shared__ int fetched;
  for_group_items(items, xofs, ctr, fetched, [&](const BatchItem &it) {
    manal_spin0_tile<T, NCB, CT, W>(job_args(jobs, it), it.major, it.minor, lds, cbuf);
  });
}

template <typename T, int NCB, int CT, int W>
__global__ void __launch_bounds__(64 * W) kb_manal_spin(const LegArgs *__restrict__ jobs,
                                                        const BatchItem *__restrict__ items,
                                                        const int *__restrict__ xofs, int *__restrict__ ctr)
{
  __shared__ __align__(16) T lds[W][2][kSlice];
  __shared__ double cbuf[2][96];
  __shared__ int fetched;
  for_group_items(items, xofs, ctr, fetched, [&](const BatchItem &it) {
    manal_spin_tile<T, NCB, CT, W>(job_args(jobs, it), it.major, it.minor, lds, cbuf);
  });
}

// ===========================================================================
// launchers
// ===========================================================================
static void check_launch(const char *what)
{
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof(buf), "kernel launch failed (%s): %s", what,
             hipGetErrorString(e));
    SHP_FAIL(buf);
  }
}

void launch_msynth(const LegArgs &a, int contr, hipStream_t st)
{
  const int nsb = (a.nslots + 63) / 64;
  const dim3 grid(xcd_grid(a.nmb, nsb));
  const dim3 block(64, MB);
  const int cols = (a.spin == 0 ? 2 : 4) * a.ntl;
  if (a.ntl < 1 || cols > 32) SHP_FAIL("bad transform group");
  const bool two = cols > 16;
  if (!a.unit_tab) SHP_FAIL("mixed synthesis without the unit-coefficient table");
  if (a.spin == 0) {
    if (two) hipLaunchKernelGGL(k_msynth_spin0<2>, grid, block, 0, st, a, nsb);
    else hipLaunchKernelGGL(k_msynth_spin0<1>, grid, block, 0, st, a, nsb);
  } else if (contr == kContrBf16x3) {
    if (!a.bsplit) SHP_FAIL("BF16 spin synthesis without the split a_lm");
    // spin 0 keeps the FP32 matrix cores: with 16 columns the contraction is
    // too small a share for the splits to pay
    if (a.deriv1) {
      if (two) hipLaunchKernelGGL((k_msynth_spin<2, true, 1>), grid, block, 0, st, a, nsb);
      else hipLaunchKernelGGL((k_msynth_spin<1, true, 1>), grid, block, 0, st, a, nsb);
    } else {
      if (two) hipLaunchKernelGGL((k_msynth_spin<2, false, 1>), grid, block, 0, st, a, nsb);
      else hipLaunchKernelGGL((k_msynth_spin<1, false, 1>), grid, block, 0, st, a, nsb);
    }
  } else if (a.deriv1) {
    if (two) hipLaunchKernelGGL((k_msynth_spin<2, true, 0>), grid, block, 0, st, a, nsb);
    else hipLaunchKernelGGL((k_msynth_spin<1, true, 0>), grid, block, 0, st, a, nsb);
  } else {
    if (two) hipLaunchKernelGGL((k_msynth_spin<2, false, 0>), grid, block, 0, st, a, nsb);
    else hipLaunchKernelGGL((k_msynth_spin<1, false, 0>), grid, block, 0, st, a, nsb);
  }
  check_launch("mixed synthesis");
}

template <int W>
static void launch_batch_anal(int kind, bool two, int contr, const LegArgs *jobs,
                              const BatchItem *items, const int *xofs, int *ctr, int grid,
                              hipStream_t st)
{
  const dim3 b(64 * W);
  auto go = [&](auto kernel) { launch_persistent(kernel, b, grid, st, jobs, items, xofs, ctr); };
  if (contr == kContrF64) {
    if (kind == kBatchAnal0) go(kb_manal_spin0<double, 1, 0, W>);
    else go(kb_manal_spin<double, 1, 0, W>);
    return;
  }
  const bool bf = contr == kContrBf16x3;
  if (kind == kBatchAnal0) {
    if (bf) { if (two) go(kb_manal_spin0<float, 2, 1, W>); else go(kb_manal_spin0<float, 1, 1, W>); }
    else { if (two) go(kb_manal_spin0<float, 2, 0, W>); else go(kb_manal_spin0<float, 1, 0, W>); }
  } else {
    if (bf) { if (two) go(kb_manal_spin<float, 2, 1, W>); else go(kb_manal_spin<float, 1, 1, W>); }
    else { if (two) go(kb_manal_spin<float, 2, 0, W>); else go(kb_manal_spin<float, 1, 0, W>); }
  }
}

template <bool FLAT>
static void batch_msynth(int kind, bool two, bool bf, dim3 b, int grid, int flat_grid,
                         hipStream_t st, const LegArgs *jobs, const BatchItem *items,
                         const int *xofs, int *ctr)
{
  auto go = [&](auto kernel) {
    launch_batch_kernel(kernel, FLAT, b, grid, flat_grid, st, jobs, items, xofs, ctr);
  };
  if (kind == kBatchSynth0) {
    // spin 0 keeps the FP32 matrix cores in both mixed forms
    if (two) go(kb_msynth_spin0<2, FLAT>); else go(kb_msynth_spin0<1, FLAT>);
  } else if (kind == kBatchSynthS) {
    if (bf) { if (two) go(kb_msynth_spin<2, false, 1, FLAT>); else go(kb_msynth_spin<1, false, 1, FLAT>); }
    else { if (two) go(kb_msynth_spin<2, false, 0, FLAT>); else go(kb_msynth_spin<1, false, 0, FLAT>); }
  } else {
    if (bf) { if (two) go(kb_msynth_spin<2, true, 1, FLAT>); else go(kb_msynth_spin<1, true, 1, FLAT>); }
    else { if (two) go(kb_msynth_spin<2, true, 0, FLAT>); else go(kb_msynth_spin<1, true, 0, FLAT>); }
  }
}

void launch_batch(const BatchLaunch &L, const LegArgs *jobs, hipStream_t st)
{
  if (L.grid <= 0) return;
  if (L.grid % 8) SHP_FAIL("persistent grid must be a multiple of 8");
  const bool two = L.ncb == 2;
  const int kind = L.kind, shape = L.shape;
  const bool synth = kind == kBatchSynth0 || kind == kBatchSynthS || kind == kBatchSynthD;
  if (synth) {
    // shape: wavefronts per workgroup, one order m each
    if (shape != 4 && shape != 8) SHP_FAIL("bad batch synthesis workgroup");
    if (L.contr == kContrF64) {
      // tg > 0: the vector-ALU kernels specialised by transform group;
      // tg == 0: the matrix-core kernels
      if (L.tg > 0)
        launch_batch_synth_f64(kind, L.tg, shape, L.flat, jobs, L.items, L.xofs, L.ctr, L.grid,
                               L.flat_grid, st);
      else
        launch_batch_synth64(kind, L.ncb, shape, L.flat, jobs, L.items, L.xofs, L.ctr, L.grid,
                             L.flat_grid, st);
      return;
    }
    const dim3 b(64, shape);
    const bool bf = L.contr == kContrBf16x3;
    if (L.flat) batch_msynth<true>(kind, two, bf, b, L.grid, L.flat_grid, st, jobs, L.items, L.xofs, L.ctr);
    else batch_msynth<false>(kind, two, bf, b, L.grid, L.flat_grid, st, jobs, L.items, L.xofs, L.ctr);
  } else {
    // shape: wavefronts (of 64 ring pairs) per workgroup
    if (shape == 4)
      launch_batch_anal<4>(kind, two, L.contr, jobs, L.items, L.xofs, L.ctr, L.grid, st);
    else if (shape == 2)
      launch_batch_anal<2>(kind, two, L.contr, jobs, L.items, L.xofs, L.ctr, L.grid, st);
    else SHP_FAIL("bad batch ring block");
  }
  check_launch("batched Legendre stage");
}

void launch_manal(const LegArgs &a, int contr, hipStream_t st)
{
  const bool f64 = contr == kContrF64, bf = contr == kContrBf16x3;
  if (a.m0 % 8 || a.m1 <= a.m0 || a.m1 > a.mmax + 1) SHP_FAIL("bad m range of an analysis launch");
  const dim3 grid(xcd_grid_mblock(a.m1 - a.m0, a.ngroups));
  const dim3 block(64 * (a.spin == 0 ? kMixWaves0 : kMixWavesS));
  const int cols = (a.spin == 0 ? 2 : 4) * a.ntl;
  // FP64: spin 0 up to 16 columns; spin s up to 32 (the kernel runs at one
  // workgroup per CU for its LDS anyway, so the second column block costs
  // no occupancy and saves a second pass of the recurrence)
  if (a.ntl < 1 || cols > ((f64 && a.spin == 0) ? 16 : 32)) SHP_FAIL("bad transform group");
  const bool two = cols > 16;
  if (f64) {
    if (a.spin == 0) hipLaunchKernelGGL((k_manal_spin0<double, 1, 0>), grid, block, 0, st, a);
    else if (two) hipLaunchKernelGGL((k_manal_spin<double, 2, 0>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_manal_spin<double, 1, 0>), grid, block, 0, st, a);
  } else if (!a.unit_tab) {
    SHP_FAIL("mixed analysis without the unit-coefficient table");
  } else if (bf) {
    if (a.spin == 0) {
      if (two) hipLaunchKernelGGL((k_manal_spin0<float, 2, 1>), grid, block, 0, st, a);
      else hipLaunchKernelGGL((k_manal_spin0<float, 1, 1>), grid, block, 0, st, a);
    } else {
      if (two) hipLaunchKernelGGL((k_manal_spin<float, 2, 1>), grid, block, 0, st, a);
      else hipLaunchKernelGGL((k_manal_spin<float, 1, 1>), grid, block, 0, st, a);
    }
  } else if (a.spin == 0) {
    if (two) hipLaunchKernelGGL((k_manal_spin0<float, 2, 0>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_manal_spin0<float, 1, 0>), grid, block, 0, st, a);
  } else {
    if (two) hipLaunchKernelGGL((k_manal_spin<float, 2, 0>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_manal_spin<float, 1, 0>), grid, block, 0, st, a);
  }
  check_launch(f64 ? "FP64 matrix-core analysis" : "mixed analysis");
}

} // namespace shp

// legendre_common.hpp -- device helpers shared by the FP64 and the mixed-
// precision Legendre kernels: exponent-scaled start values and the spin-0 /
// spin-s recurrence steps, in libsharp's operation order
// (libsharp/sharp_core_inc.c:73-293).
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include "shp_internal.hpp"
#include "kernels.hpp"

namespace shp {

// Maps a linear workgroup id to (major, minor) so that the nminor workgroups
// of one major index land on the same XCD (the hardware deals workgroup ids
// round-robin over the 8 XCDs, each with its own L2): workgroups that read
// the same rows share one L2.  False for the padding ids of the last round.
__device__ inline bool xcd_map(int id, int nmajor, int nminor, int &major, int &minor)
{
  const int xcd = id & 7, k = id >> 3;
  major = (k / nminor) * 8 + xcd;
  minor = k % nminor;
  return major < nmajor;
}

static inline int xcd_grid(int nmajor, int nminor)
{
  return ((nmajor + 7) / 8) * 8 * nminor;
}

// Wave-uniform rows (recurrence coefficients, canonical a_lm) read through
// the constant address space.  The rows do not change during a kernel; with
// that known the compiler turns a wave-uniform address into a scalar load
// even where the row pointer itself was read from memory (the job arguments
// of a batch) -- for a pointer that is a kernel argument it finds this out by
// itself -- and the values stay in scalar registers.
typedef const __attribute__((address_space(4))) double *ConstRowPtr;
struct URow {
  ConstRowPtr q;
  __device__ explicit URow(const double *p) : q((ConstRowPtr)p) {}
  __device__ double operator[](int64_t i) const { return q[i]; }
};
struct URow2 {
  ConstRowPtr q;
  __device__ explicit URow2(const double2 *p) : q((ConstRowPtr)p) {}
  __device__ double2 operator[](int64_t i) const { return make_double2(q[2 * i], q[2 * i + 1]); }
};

// Wave-uniform copy of a struct read from memory: each 32-bit word through
// readfirstlane, so that the compiler keeps the fields in scalar registers
// (as it does for kernel arguments) instead of vector registers, where the
// job arguments of a batch tile cost about 40 registers per lane and with
// them resident wavefronts.
template <typename T> __device__ inline T uniform_load(const T *p)
{
  static_assert(sizeof(T) % 4 == 0, "whole words");
  T v;
  const int *src = reinterpret_cast<const int *>(p);
  int *dst = reinterpret_cast<int *>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return v;
}

// Batch kernels.  The work list is split into 8 contiguous parts, xofs[0..8],
// one per XCD (the host groups the items that share rows into one part);
// workgroups with blockIdx.x % 8 == x, which the hardware places on XCD x,
// take the items of part x.  Every item is a tile of one workgroup.
//
// Flat form: workgroup b takes item b/8 of part b%8 (grid: 8 x the largest
// part); the hardware deals the workgroups as resources free up.
template <typename F>
__device__ inline void for_flat_item(const BatchItem *items, const int *xofs, F &&f)
{
  const int x = blockIdx.x & 7, p = xofs[x] + (blockIdx.x >> 3);
  if (p < xofs[x + 1]) f(uniform_load(items + p));
}

// Persistent form: each workgroup takes items of its part from the counter
// ctr[x] (zeroed before the launch) until the part is used up, so that
// workgroups that start late, or are slow, simply take fewer items.  The
// first thread fetches, the workgroup follows.
template <typename F>
__device__ inline void for_group_items(const BatchItem *items, const int *xofs, int *ctr,
                                       int &fetched, F &&f)
{
  const int x = blockIdx.x & 7, lo = xofs[x], n = xofs[x + 1] - lo;
  for (;;) {
    if (threadIdx.x == 0 && threadIdx.y == 0) fetched = atomicAdd(&ctr[x], 1);
    __syncthreads();
    const int p = fetched;
    __syncthreads();   // every thread has read `fetched` before the next fetch
    if (p >= n) break;
    f(uniform_load(items + lo + p));
    __syncthreads();   // the next item reuses the LDS of this one
  }
}

// zeroes the 8 per-XCD item counters of a persistent launch (a kernel, not
// a memset, so that graphs replay it correctly)
template <int N> __global__ void k_zero_ints(int *p)
{
  if (threadIdx.x < N) p[threadIdx.x] = 0;
}

// Launches a batch kernel: flat (grid flat_grid, no counters) or persistent
// (below).
template <typename K>
static inline void launch_batch_kernel(K kernel, bool flat, dim3 block, int grid, int flat_grid,
                                       hipStream_t st, const LegArgs *jobs,
                                       const BatchItem *items, const int *xofs, int *ctr);

// Launches a persistent batch kernel: its counters zeroed first on the same
// stream, and its grid (a multiple of 8) no larger than the workgroups the
// device keeps resident at once (more would only find the work taken).
template <typename K>
static inline void launch_persistent(K kernel, dim3 block, int grid, hipStream_t st,
                                     const LegArgs *jobs, const BatchItem *items,
                                     const int *xofs, int *ctr)
{
  hipLaunchKernelGGL(k_zero_ints<8>, dim3(1), dim3(64), 0, st, ctr);
  int nb = 0, dev = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block.x * block.y * block.z, 0) !=
          hipSuccess || nb < 1)
    nb = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 256;
  const int cap = std::max(8, (nb * ncu / 8) * 8);
  hipLaunchKernelGGL(kernel, dim3(std::min(grid, cap)), block, 0, st, jobs, items, xofs, ctr);
}

template <typename K>
static inline void launch_batch_kernel(K kernel, bool flat, dim3 block, int grid, int flat_grid,
                                       hipStream_t st, const LegArgs *jobs,
                                       const BatchItem *items, const int *xofs, int *ctr)
{
  if (flat) {
    if (flat_grid > 0)
      hipLaunchKernelGGL(kernel, dim3(flat_grid), block, 0, st, jobs, items, xofs, ctr);
  } else {
    launch_persistent(kernel, block, grid, st, jobs, items, xofs, ctr);
  }
}

// the launch arguments of an item's job, for its transform group
__device__ inline LegArgs job_args(const LegArgs *jobs, const BatchItem &it)
{
  LegArgs a = uniform_load(jobs + it.job);
  a.t0 = it.t0;
  a.ntl = it.ntl;
  a.part_col0 = it.col0;
  return a;
}

// Brings |val| into [maxval*2^-800, maxval].  A finite double needs at most
// three 800-binade steps; the bound also ends the loop for non-finite input.
__device__ inline void normalize(double &val, int &scale, double maxval)
{
  for (int it = 0; it < 4 && fabs(val) > maxval; ++it) { val *= kFSmall; ++scale; }
  if (val != 0.)
    for (int it = 0; it < 4 && fabs(val) < maxval * kFSmall; ++it) {
      val *= kFBig;
      --scale;
    }
}

// val^npow with a separate exponent (libsharp/sharp_core_inc.c:97-126)
__device__ inline void mypow(double val, int npow, double &resd, int &ress)
{
  int scale = 0, scaleint = 0;
  double res = 1.;
  normalize(val, scaleint, kFBigHalf);
  do {
    if (npow & 1) {
      res *= val;
      scale += scaleint;
      normalize(res, scale, kFBigHalf);
    }
    val *= val;
    scaleint += scaleint;
    normalize(val, scaleint, kFBigHalf);
  } while (npow >>= 1);
  resd = res;
  ress = scale;
}

__device__ inline double corfac(int scale)
{
  return (scale < 0) ? 0. : ((scale == 0) ? 1. : kFBig);
}

__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }
__device__ inline bool wave_all(bool p) { return __ballot(!p) == 0ull; }

struct SpinStart {
  int mhi, cosPow, sinPow, preMinus_p, preMinus_m;
};

__device__ inline SpinStart spin_start(int m, int s)
{
  SpinStart r;
  const int mhi = (m < s) ? s : m;
  r.mhi = mhi;
  r.preMinus_p = r.preMinus_m = 0;
  if (mhi == m) {
    r.cosPow = mhi + s;
    r.sinPow = mhi - s;
    r.preMinus_p = r.preMinus_m = ((mhi - s) & 1);
  } else {
    r.cosPow = mhi + m;
    r.sinPow = mhi - m;
    r.preMinus_m = ((mhi + m) & 1);
  }
  return r;
}

// start values of lambda_plus / lambda_minus at l=mhi with their scales
// (libsharp/sharp_core_inc.c:223-273)
__device__ inline void spin_init(double cth, double sth, int s, const SpinStart &st,
                                 double prefac, int fscale, double &rec2p,
                                 double &rec2m, int &scalep, int &scalem)
{
  double cth2 = fmax(sqrt((1. + cth) * 0.5), 1e-15);
  double sth2 = fmax(sqrt((1. - cth) * 0.5), 1e-15);
  if (sth < 0. && cth < 0.) cth2 = -cth2;
  if (sth < 0. && cth > 0.) sth2 = -sth2;
  double ccp, ssp, csp, scp;
  int ccps, ssps, csps, scps;
  mypow(cth2, st.cosPow, ccp, ccps);
  mypow(sth2, st.sinPow, ssp, ssps);
  mypow(cth2, st.sinPow, csp, csps);
  mypow(sth2, st.cosPow, scp, scps);
  rec2p = prefac * ccp; scalep = fscale + ccps;
  rec2m = prefac * csp; scalem = fscale + csps;
  normalize(rec2m, scalem, kFBigHalf);
  normalize(rec2p, scalep, kFBigHalf);
  rec2p *= ssp; scalep += ssps;
  rec2m *= scp; scalem += scps;
  if (st.preMinus_p) rec2p = -rec2p;
  if (st.preMinus_m) rec2m = -rec2m;
  if (s & 1) rec2p = -rec2p;
  normalize(rec2m, scalem, kFTol);
  normalize(rec2p, scalep, kFTol);
}

// spin-s recurrence coefficients of one l: libsharp's fx[l] = {f0, f1, f2}
// (libsharp/sharp_ylmgen_c.c:174-183) with f1 stored premultiplied, g = f0*f1
struct Fx { double f0, g, f2; };

// x_{l+1} = (cth -+ f1) f0 y_l - f2 x_{l-1} (libsharp/sharp_core_inc.c:
// 209-221), the factor (cth -+ f1) f0 = cth f0 -+ g as one FMA
__device__ inline void rec_step(double &rxp, double &rxm, double ryp, double rym,
                                double cth, const Fx &f)
{
  rxp = fma(fma(cth, f.f0, -f.g), ryp, -(f.f2 * rxp));
  rxm = fma(fma(cth, f.f0, f.g), rym, -(f.f2 * rxm));
}

// rec_step on the table whose f2 are all 1: the recurrence of lambda_l / s_l
// with s_l = f2_l s_{l-2} (s = 1 at the first two l), whose factor s_l the
// a_lm gather and scatter apply instead (the mixed-precision kernels)
__device__ inline void rec_step_unit(double &rxp, double &rxm, double ryp, double rym,
                                     double cth, const Fx &f)
{
  rxp = fma(fma(cth, f.f0, -f.g), ryp, -rxp);
  rxm = fma(fma(cth, f.f0, f.g), rym, -rxm);
}

__device__ inline bool rescale1(double &l1, double &l2, int &scale)
{
  if (fabs(l2) > kFTol) {
    l1 *= kFSmall; l2 *= kFSmall; ++scale;
    return true;
  }
  return false;
}

} // namespace shp

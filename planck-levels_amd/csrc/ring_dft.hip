// ring_dft.hip -- ring stage by direct summation, and a_lm layout kernels.
//
// The ring stage evaluates, for every ring of nph pixels at azimuths
// phi_j = phi0 + 2 pi j / nph,
//   synthesis  f_j = Re c_0 + 2 sum_{m=1}^{mmax} Re(c_m e^{i m phi_j}),
//              c_m = w * phase_m,
//   analysis   phase_m = w e^{-i m phi0} sum_j f_j e^{-2 pi i m j / nph},
// which are exactly the sums libsharp forms with its aliasing fold and real
// FFTs (libsharp/sharp.c:253-353).  This direct form costs O(nph * mmax) per
// ring; it is exact for every ring length and serves as the reference path
// of the batched ring FFT in ring_fft.hip.  Twiddles are anchored every 32
// steps on an exactly range-reduced angle (integer (m*j) mod nph) and advanced
// by complex multiplication in between.
//
// The a_lm kernels move coefficients between the caller's layout
// (mvstart/stride, SHARP_PACKED, float or double) and the canonical almtmp,
// applying norm_l and the REAL_HARMONICS m=0 factor
// (libsharp/sharp.c:481-606).

#include <algorithm>
#include <hip/hip_runtime.h>
#include <type_traits>
#include "alm_finish.hpp"
#include "kernels.hpp"
#include "sharp_lowlevel.h"

namespace shp {

constexpr double kSqrt2 = 1.414213562373095145474621858739;

// e^{2 pi i k / n} for 0 <= k < n
__device__ inline double2 unit_root(int64_t k, int n)
{
  double s, c;
  sincospi(2.0 * (double)k / (double)n, &s, &c);
  return make_double2(c, s);
}

__device__ inline double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---------------------------------------------------------------------------
// phase -> ring, direct summation.  grid: (2*nslots, ncomp, pixel tiles)
// ---------------------------------------------------------------------------
constexpr int kTileM = 512;

template <typename T>
__global__ void __launch_bounds__(256) k_phase2ring_dft(RingArgs a, const int *list)
{
  __shared__ double2 cm[kTileM];
  const int rid = list[blockIdx.x];
  const int slot = rid >> 1, hemi = rid & 1, comp = blockIdx.y;
  const RingDev r = a.ring[rid];
  const int n = r.nph;
  if (n <= 0) return;
  const int j = blockIdx.z * 256 + threadIdx.x;
  if ((int)(blockIdx.z * 256) >= n) return;
  double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.weight : 1.;
  if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrtHalf;
  const bool norot = fabs(r.phi0) < 1e-14;
  double acc = 0.;
  const double2 step = unit_root(j < n ? j : 0, n);
  for (int m0 = 0; m0 <= a.mmax; m0 += kTileM) {
    __syncthreads();
    for (int i = threadIdx.x; i < kTileM; i += 256) {
      const int m = m0 + i;
      double2 v = make_double2(0., 0.);
      if (m <= a.mmax) {
        v = a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)];
        v.x *= wgt; v.y *= wgt;
        if (!norot && m > 0) {
          double s, c;
          sincos(m * r.phi0, &s, &c);
          v = cmul(v, make_double2(c, s));
        }
        if (m > 0) { v.x *= 2.; v.y *= 2.; }
        else v.y = 0.;
      }
      cm[i] = v;
    }
    __syncthreads();
    if (j < n) {
      const int mend = min(kTileM, a.mmax + 1 - m0);
      for (int i0 = 0; i0 < mend; i0 += 32) {
        double2 tw = unit_root(((int64_t)(m0 + i0) * j) % n, n);
        const int iend = min(i0 + 32, mend);
        for (int i = i0; i < iend; ++i) {
          acc += cm[i].x * tw.x - cm[i].y * tw.y;
          tw = cmul(tw, step);
        }
      }
    }
  }
  if (j < n) {
    const int64_t idx = r.ofs + (int64_t)j * r.stride - a.map_base;
    void *mp = a.maps.p[comp];
    const double old = (a.flags & SHARP_ADD) ? (a.dp ? ld<double>(mp, idx) : ld<float>(mp, idx)) : 0.;
    if (a.dp) st<double>(mp, idx, old + acc);
    else st<float>(mp, idx, old + acc);
  }
}

// ---------------------------------------------------------------------------
// ring -> phase, direct summation.  grid: (2*nslots, ncomp, m tiles)
// ---------------------------------------------------------------------------
constexpr int kTileJ = 1024;

__global__ void __launch_bounds__(256) k_ring2phase_dft(RingArgs a, const int *list)
{
  __shared__ double f[kTileJ];
  const int rid = list[blockIdx.x];
  const int slot = rid >> 1, hemi = rid & 1, comp = blockIdx.y;
  const RingDev r = a.ring[rid];
  const int n = r.nph;
  if (n <= 0) return;
  const int m = blockIdx.z * 256 + threadIdx.x;
  double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.weight : 1.;
  if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrt2;
  const void *mp = a.maps.p[comp];
  double2 acc = make_double2(0., 0.);
  // e^{-2 pi i m / n}
  double2 step = unit_root((int64_t)(n - (m % n)) % n, n);
  for (int j0 = 0; j0 < n; j0 += kTileJ) {
    __syncthreads();
    for (int i = threadIdx.x; i < kTileJ; i += 256) {
      const int jj = j0 + i;
      double v = 0.;
      if (jj < n) {
        const int64_t idx = r.ofs + (int64_t)jj * r.stride - a.map_base;
        v = a.dp ? ld<double>(mp, idx) : ld<float>(mp, idx);
      }
      f[i] = v;
    }
    __syncthreads();
    if (m <= a.mmax) {
      const int jend = min(kTileJ, n - j0);
      for (int i0 = 0; i0 < jend; i0 += 32) {
        const int64_t k = ((int64_t)m * (j0 + i0)) % n;
        double2 tw = unit_root((n - k) % n, n);
        const int iend = min(i0 + 32, jend);
        for (int i = i0; i < iend; ++i) {
          acc.x += f[i] * tw.x;
          acc.y += f[i] * tw.y;
          tw = cmul(tw, step);
        }
      }
    }
  }
  if (m > a.mmax) return;
  acc.x *= wgt; acc.y *= wgt;
  if (fabs(r.phi0) >= 1e-14) {
    double s, c;
    sincos(-m * r.phi0, &s, &c);
    acc = cmul(acc, make_double2(c, s));
  }
  a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)] = acc;
}

// ---------------------------------------------------------------------------
// SHARP_NO_FFT: the map holds nph = mmax+1 complex values per ring
// (libsharp/sharp.c:631-669).  grid: (2*nslots, ncomp)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_phase2ring_direct(RingArgs a, const int *list)
{
  const int rid = list[blockIdx.x];
  const int slot = rid >> 1, hemi = rid & 1, comp = blockIdx.y;
  const RingDev r = a.ring[rid];
  if (r.nph <= 0) return;
  double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.nph * r.weight : 1.;
  if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrtHalf;
  void *mp = a.maps.p[comp];
  for (int m = threadIdx.x; m <= a.mmax; m += 256) {
    const double2 v = a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)];
    const int64_t idx = 2 * (r.ofs + (int64_t)m * r.stride - a.map_base);
    double ore = 0., oim = 0.;
    if (a.flags & SHARP_ADD) {
      ore = a.dp ? ld<double>(mp, idx) : ld<float>(mp, idx);
      oim = a.dp ? ld<double>(mp, idx + 1) : ld<float>(mp, idx + 1);
    }
    if (a.dp) { st<double>(mp, idx, ore + wgt * v.x); st<double>(mp, idx + 1, oim + wgt * v.y); }
    else { st<float>(mp, idx, ore + wgt * v.x); st<float>(mp, idx + 1, oim + wgt * v.y); }
  }
}

__global__ void __launch_bounds__(256) k_ring2phase_direct(RingArgs a, const int *list)
{
  const int rid = list[blockIdx.x];
  const int slot = rid >> 1, hemi = rid & 1, comp = blockIdx.y;
  const RingDev r = a.ring[rid];
  if (r.nph <= 0) return;
  double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.nph * r.weight : 1.;
  if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrt2;
  const void *mp = a.maps.p[comp];
  for (int m = threadIdx.x; m <= a.mmax; m += 256) {
    const int64_t idx = 2 * (r.ofs + (int64_t)m * r.stride - a.map_base);
    const double re = a.dp ? ld<double>(mp, idx) : ld<float>(mp, idx);
    const double im = a.dp ? ld<double>(mp, idx + 1) : ld<float>(mp, idx + 1);
    a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)] =
        make_double2(re * wgt, im * wgt);
  }
}

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

void launch_phase2ring_dft(const RingArgs &a, const int *list, int nlist, hipStream_t s)
{
  if (nlist <= 0) return;
  if (a.flags & SHARP_NO_FFT) {
    hipLaunchKernelGGL(k_phase2ring_direct, dim3(nlist, a.ncomp), dim3(256), 0, s, a, list);
  } else {
    const dim3 grid(nlist, a.ncomp, (a.nphmax + 255) / 256);
    if (a.dp) hipLaunchKernelGGL(k_phase2ring_dft<double>, grid, dim3(256), 0, s, a, list);
    else hipLaunchKernelGGL(k_phase2ring_dft<float>, grid, dim3(256), 0, s, a, list);
  }
  check_launch("phase2ring");
}

void launch_ring2phase_dft(const RingArgs &a, const int *list, int nlist, hipStream_t s)
{
  if (nlist <= 0) return;
  if (a.flags & SHARP_NO_FFT) {
    hipLaunchKernelGGL(k_ring2phase_direct, dim3(nlist, a.ncomp), dim3(256), 0, s, a, list);
  } else {
    const dim3 grid(nlist, a.ncomp, (a.mmax + 256) / 256);
    hipLaunchKernelGGL(k_ring2phase_dft, grid, dim3(256), 0, s, a, list);
  }
  check_launch("ring2phase");
}

// ---------------------------------------------------------------------------
// a_lm gather / scatter.  grid: (tiles of (l, component) pairs of one m, m);
// element e of row m is l = m + e / ncomp, component e % ncomp, so the
// canonical rows are read and written contiguously.
// ---------------------------------------------------------------------------
// a_{l,m} of component comp from the caller's array, times norm_l (and sqrt 2
// for REAL_HARMONICS at m = 0); only the real part for m = 0
// (libsharp/sharp.c:481-533)
template <typename T>
__device__ inline double2 gather_value(const AlmArgs &a, int m, int l, int comp)
{
  const bool realonly = a.packed && m == 0;
  int64_t ofs = a.mvstart_by_m[m], stride = a.stride;
  if (!a.packed) ofs *= 2;
  if (!realonly) stride *= 2;
  const int64_t idx = ofs + (int64_t)l * stride - a.alm_base;
  const void *p = a.alms.p[comp];
  const double re = ld<T>(p, idx), im = (realonly || m == 0) ? 0. : ld<T>(p, idx + 1);
  double f = a.norm_l[l];
  if (a.lsc) f *= a.lsc[tri_ofs(m, a.lmax) + l - m];
  if (m == 0 && (a.flags & SHARP_REAL_HARMONICS)) f *= kSqrt2;
  return make_double2(re * f, im * f);
}

// The m of canonical row `row` (rows of m at tri_ofs(m, lmax) ..): the
// root of m (2 lmax + 3 - m) / 2 = row, then an exact integer correction
__device__ inline int row_m(int64_t row, int lmax)
{
  const double b = 2.0 * lmax + 3.0;
  int m = (int)((b - sqrt(fmax(b * b - 8.0 * (double)row, 0.))) * 0.5);
  while (m > 0 && tri_ofs(m, lmax) > row) --m;
  while (tri_ofs(m + 1, lmax) <= row) ++m;
  return m;
}

// one thread per (row, component) of the triangle, as the scatter
template <typename T>
__global__ void __launch_bounds__(256) k_alm_gather(AlmArgs a)
{
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int nc = a.ncomp;
  if (e >= tri_ofs(a.mmax + 1, a.lmax) * nc) return;
  const int64_t row = e / nc;
  const int comp = (int)(e - row * nc);
  const int m = row_m(row, a.lmax);
  const int l = m + (int)(row - tri_ofs(m, a.lmax));
  const double2 v = gather_value<T>(a, m, l, comp);
  const int64_t o = e;
  if (a.tmp_f32)
    reinterpret_cast<float2 *>(a.almtmp)[o] = make_float2((float)v.x, (float)v.y);
  else
    a.almtmp[o] = v;
}

// Gather of a BF16 spin-s synthesis: the FP32 canonical rows as k_alm_gather
// writes them, and the BF16 piece words of the output columns (bs_pairs).
// One workgroup per (m, 64 row pairs from o = max(m, spin), 16 components)
// goes through an LDS tile, so that the reads of the caller's a_lm (along l),
// the FP32 rows (along the columns) and the piece words (along the pairs) are
// all coalesced; 16 components are 8 whole transforms (E, B) or 16 (E only),
// so the columns c and c^3 of a word are in one tile.  The rows m <= l < o
// (m < spin) get FP32 rows only; the last output column is the zero column.
template <typename T>
__global__ void __launch_bounds__(256) k_alm_gather_split(AlmArgs a, int ntiles)
{
  __shared__ float tile[128][33];
  const int m = blockIdx.y;
  const int o0 = m < a.spin ? a.spin : m;
  const int np = bs_pairs(a.lmax, o0);
  const int tix = blockIdx.x % ntiles, chunk = blockIdx.x / ntiles;
  const int r0 = tix * 64, c0 = chunk * 16, nc = a.ncomp, ncf = 2 * nc;
  float *rows = reinterpret_cast<float *>(a.almtmp) + ((int64_t)tri_ofs(m, a.lmax) - m) * ncf;
  if (tix == 0) {
    for (int e = threadIdx.x; e < (o0 - m) * 16; e += 256) {
      const int l = m + e / 16, comp = c0 + e % 16;
      if (l > a.lmax || comp >= nc) continue;
      const double2 v = gather_value<T>(a, m, l, comp);
      rows[(int64_t)l * ncf + 2 * comp] = (float)v.x;
      rows[(int64_t)l * ncf + 2 * comp + 1] = (float)v.y;
    }
  }
  if (r0 >= np) return;
  const int nr = min(64, np - r0), ncomp = min(16, nc - c0);
  const int lb = o0 + 2 * r0;
  {
    // a thread's tile row i (one l) and its components cc = tid/128 + 2u:
    // the scale and offset once (as gather_value forms them), then the
    // loads of all 8 components in flight together
    const int i = threadIdx.x & 127, l = lb + i;
    const bool lok = i < 2 * nr && l <= a.lmax;
    double f = 0.;
    int64_t idx = 0;
    if (lok) {
      const bool realonly = a.packed && m == 0;
      f = a.norm_l[l];
      if (a.lsc) f *= a.lsc[tri_ofs(m, a.lmax) + l - m];
      if (m == 0 && (a.flags & SHARP_REAL_HARMONICS)) f *= kSqrt2;
      int64_t ofs = a.mvstart_by_m[m], stride = a.stride;
      if (!a.packed) ofs *= 2;
      if (!realonly) stride *= 2;
      idx = ofs + (int64_t)l * stride - a.alm_base;
    }
    double re[8], im[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int comp = c0 + (threadIdx.x >> 7) + 2 * u;
      re[u] = 0.;
      im[u] = 0.;
      if (lok && comp < nc) {
        const void *p = a.alms.p[comp];
        re[u] = ld<T>(p, idx);
        if (m != 0) im[u] = ld<T>(p, idx + 1);
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cc = (threadIdx.x >> 7) + 2 * u;
      tile[i][2 * cc] = (float)(re[u] * f);
      tile[i][2 * cc + 1] = (float)(im[u] * f);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 128 * 32; e += 256) {
    const int i = e >> 5, cc = e & 31, l = lb + i, c = 2 * c0 + cc;
    if (i < 2 * nr && l <= a.lmax && c < ncf) rows[(int64_t)l * ncf + c] = tile[i][cc];
  }
  unsigned *w = a.bsplit + a.bs_base[m];
  // output columns of this chunk: 4 per transform, from column ocb on
  const int nout = a.bs_deriv ? 4 * ncomp : 2 * ncomp;
  const int ocb = a.bs_deriv ? 4 * c0 : 2 * c0;
  // X of output column k (tile row i), or 0
  auto xval = [&](int i, int k) {
    if (!a.bs_deriv) return tile[i][k];
    return (k & 3) < 2 ? tile[i][2 * (k >> 2) + (k & 3)] : 0.f;
  };
  for (int e = threadIdx.x; e < nout * 3 * 64; e += 256) {
    const int p = e & 63, q = (e >> 6) % 3, k = (e >> 6) / 3;
    if (p >= nr) continue;
    // piece q of the exact three-piece split (x = h + m + l, truncation);
    // the split of -x is the negated split of x
    auto piece = [&](float x) {
      const unsigned xu = __float_as_uint(x);
      if (q == 0) return xu >> 16;
      const float r1 = x - __uint_as_float(xu & 0xffff0000u);
      const unsigned ru = __float_as_uint(r1);
      if (q == 1) return ru >> 16;
      return __float_as_uint(r1 - __uint_as_float(ru & 0xffff0000u)) >> 16;
    };
    const int kk = k & 3;
    const float sy = (kk == 0 || kk == 3) ? -1.f : 1.f;
    const float x = xval(2 * p, k), y = sy * xval(2 * p + 1, k ^ 3);
    w[(int64_t)(3 * (ocb + k) + q) * np + r0 + p] = (piece(y) << 16) | piece(x);
  }
  if (c0 + 16 >= nc) {
    const int zc = a.bs_ncol - 1;
    for (int e = threadIdx.x; e < 3 * 64; e += 256) {
      const int p = e & 63, q = e >> 6;
      if (p < nr) w[(int64_t)(3 * zc + q) * np + r0 + p] = 0u;
    }
  }
}

void launch_alm_gather_split(const AlmArgs &a, hipStream_t s)
{
  // tiles from the longest m (pairs from max(m, spin)); chunks of 16
  // components, the last one also writing the zero column
  const int ntiles = std::max(1, (bs_pairs(a.lmax, a.spin) + 63) / 64);
  const int nchunks = std::max(1, (a.ncomp + 15) / 16);
  const int nout = a.bs_deriv ? 4 * a.ncomp : 2 * a.ncomp;
  if (a.bs_ncol != nout + 1) SHP_FAIL("BF16 a_lm split: bad column count");
  if (!a.bs_deriv && (a.ncomp & 1)) SHP_FAIL("BF16 a_lm split: E and B per transform expected");
  const dim3 grid(ntiles * nchunks, a.mmax + 1);
  if (a.flags & SHARP_DP)
    hipLaunchKernelGGL(k_alm_gather_split<double>, grid, dim3(256), 0, s, a, ntiles);
  else
    hipLaunchKernelGGL(k_alm_gather_split<float>, grid, dim3(256), 0, s, a, ntiles);
  check_launch("alm_gather_split");
}

// components [c0, c0+npc); the values come from almtmp, or (part != nullptr)
// from the partial sums of the ring-pair groups that reach m, added in FP64
// in group order
// one thread per (row, component) of the triangle: a flat grid with no idle
// workgroups
template <typename T>
__global__ void __launch_bounds__(256) k_alm_scatter(AlmArgs a)
{
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int npc = a.npc;
  const int64_t r0 = tri_ofs(a.m0, a.lmax);
  if (e >= (tri_ofs(a.m1, a.lmax) - r0) * npc) return;
  const int64_t q = e / npc, row = r0 + q;
  const int cl = (int)(e - q * npc), comp = a.c0 + cl;
  const int m = row_m(row, a.lmax);
  const int l = m + (int)(row - tri_ofs(m, a.lmax));
  double2 x;
  if (a.part) {
    const GroupRange gr = group_range(a, m);
    x = a.part_f64 ? part_sum<true>(a, gr, m, row, comp) : part_sum<false>(a, gr, m, row, comp);
  } else {
    x = a.almtmp[row * a.ncomp + comp];
  }
  alm_store<T>(a, m, l, row, comp, x);
}

void launch_alm_gather(const AlmArgs &a, hipStream_t s)
{
  const int64_t tot = tri_ofs(a.mmax + 1, a.lmax) * a.ncomp;
  const dim3 grid((unsigned)((tot + 255) / 256));
  if (a.flags & SHARP_DP) hipLaunchKernelGGL(k_alm_gather<double>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_alm_gather<float>, grid, dim3(256), 0, s, a);
  check_launch("alm_gather");
}

template <typename T>
__global__ void __launch_bounds__(256) k_map_residual(PtrTab src, PtrTab dst, int64_t n)
{
  const T *x = static_cast<const T *>(src.p[blockIdx.y]);
  T *y = static_cast<T *>(dst.p[blockIdx.y]);
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * 256)
    y[k] = x[k] - y[k];
}

void launch_map_residual(const PtrTab &src, const PtrTab &dst, int ncomp, int64_t n,
                         bool dp, hipStream_t s)
{
  if (n <= 0 || ncomp <= 0) return;
  const dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 4096), ncomp);
  if (dp) hipLaunchKernelGGL(k_map_residual<double>, grid, dim3(256), 0, s, src, dst, n);
  else hipLaunchKernelGGL(k_map_residual<float>, grid, dim3(256), 0, s, src, dst, n);
  check_launch("map_residual");
}

void launch_alm_scatter(const AlmArgs &a, hipStream_t s)
{
  if (a.m0 < 0 || a.m1 <= a.m0 || a.m1 > a.mmax + 1) SHP_FAIL("bad m range of the a_lm scatter");
  const int64_t tot = (tri_ofs(a.m1, a.lmax) - tri_ofs(a.m0, a.lmax)) * a.npc;
  const dim3 grid((unsigned)((tot + 255) / 256));
  if (a.flags & SHARP_DP) hipLaunchKernelGGL(k_alm_scatter<double>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_alm_scatter<float>, grid, dim3(256), 0, s, a);
  check_launch("alm_scatter");
}

} // namespace shp

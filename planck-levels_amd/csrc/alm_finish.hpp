// alm_finish.hpp -- the last step of an analysis: the a_lm of one (row,
// component) summed from the per-group partial sums of the matrix-core
// analysis kernels, scaled and stored in the caller's layout
// (libsharp/sharp.c:536-606, almtmp2alm), as the a_lm scatter kernel does it.

#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "kernels.hpp"
#include "sharp_lowlevel.h"

namespace shp {

constexpr double kSqrtHalf = 0.707106781186547572737310929369;

template <typename T> __device__ inline double ld(const void *p, int64_t i)
{
  return (double)static_cast<const T *>(p)[i];
}
template <typename T> __device__ inline void st(void *p, int64_t i, double v)
{
  static_cast<T *>(p)[i] = (T)v;
}

// The groups that reach m (gmax[g] >= m) are added in group order.  When
// gmax does not decrease along the groups (ring pairs ordered from a pole
// towards the equator, as in every iso-latitude grid) they are the groups
// from the first one that reaches m on (glo).
struct GroupRange {
  int glo;
  bool mono;
};

__device__ inline GroupRange group_range(const AlmArgs &a, int m)
{
  GroupRange r{0, true};
  for (int g = 0; g < a.ngroups; ++g) {
    const int gm = a.gmax[g];
    r.glo += gm < m;
    if (g > 0) r.mono = r.mono && gm >= a.gmax[g - 1];
  }
  return r;
}

// sum over the groups that reach m of the partial value of (row, comp), in
// FP64, in group order; with monotone limits the loads are issued four at a
// time (indices past the last group read the last group again, unused)
template <bool F64>
__device__ inline double2 part_sum(const AlmArgs &a, GroupRange gr, int m, int64_t row, int comp)
{
  const int64_t gstride = a.part_rows * a.part_ld;
  const int64_t o = (row - a.part_row0) * a.part_ld + 2 * (comp - a.part_c0);
  auto part_at = [&](int g) {
    if constexpr (F64)
      return *reinterpret_cast<const double2 *>(static_cast<const double *>(a.part) + o +
                                                g * gstride);
    else {
      const float2 v = *reinterpret_cast<const float2 *>(static_cast<const float *>(a.part) + o +
                                                         g * gstride);
      return make_double2(v.x, v.y);
    }
  };
  double re = 0., im = 0.;
  if (gr.mono) {
    for (int g0 = gr.glo; g0 < a.ngroups; g0 += 4) {
      double2 v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = part_at(min(g0 + i, a.ngroups - 1));
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (g0 + i < a.ngroups) { re += v[i].x; im += v[i].y; }
    }
  } else {
    for (int g = 0; g < a.ngroups; ++g) {
      if (a.gmax[g] < m) continue;
      const double2 v = part_at(g);
      re += v.x;
      im += v.y;
    }
  }
  return make_double2(re, im);
}

// a_{l,m} of component comp = x norm_l (s_{m,l}) (sqrt(1/2) for
// REAL_HARMONICS at m = 0) into the caller's array, added under SHARP_ADD;
// only the real part for m = 0 (the imaginary part zeroed unless PACKED)
template <typename T>
__device__ inline void alm_store(const AlmArgs &a, int m, int l, int64_t row, int comp, double2 x)
{
  const bool realonly = a.packed && m == 0;
  int64_t ofs = a.mvstart_by_m[m], stride = a.stride;
  if (!a.packed) ofs *= 2;
  if (!realonly) stride *= 2;
  const int64_t idx = ofs + (int64_t)l * stride - a.alm_base;
  void *p = a.alms.p[comp];
  double f = a.norm_l[l];
  if (a.lsc) f *= a.lsc[row];
  if (m == 0 && (a.flags & SHARP_REAL_HARMONICS)) f *= kSqrtHalf;
  const bool add = (a.flags & SHARP_ADD) != 0;
  st<T>(p, idx, (add ? ld<T>(p, idx) : 0.) + x.x * f);
  if (realonly) return;
  if (m == 0) {
    if (!add) st<T>(p, idx + 1, 0.);
  } else {
    st<T>(p, idx + 1, (add ? ld<T>(p, idx + 1) : 0.) + x.y * f);
  }
}

} // namespace shp

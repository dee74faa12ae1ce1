// shp_internal.hpp -- internal declarations of the MI355X SHT engine.
//
// Device data layout (all FP64 unless stated):
//
//  * slots:   the ring pairs of a geometry in device order (sorted by
//             sin(theta), so that a 64-lane wavefront sees similar rings and
//             the per-ring m cutoff prunes whole wavefronts).  Per slot:
//             cth, sth, mlim, ispair and two RingDev (north r1, south r2).
//  * phase:   phase[comp][m/MB][slot][hemi][m%MB], complex; MB=8 consecutive
//             m of one ring are contiguous (128 B), so the ring transforms
//             read and write whole cache lines while the Legendre kernels,
//             which run one m per wavefront, fill each line over MB waves.
//  * almtmp:  canonical a_lm, complex, index (tri(m)+l-m)*ncomp + comp with
//             tri(m) = m*(lmax+1) - m*(m-1)/2; norm_l and the
//             REAL_HARMONICS m=0 factor already applied.
//  * tables:  per-m recurrence coefficients in rows of (lmax-m+1+kTabPad)
//             entries: spin 0 {f0,f1} (double2), spin s {f0,f1,f2} (3 doubles);
//             entries past lmax are zero so that recurrences may run a few
//             steps beyond lmax without reading outside a row.
#pragma once

#include <cstddef>
#include <cstdint>

#define SHP_FAIL(msg) ::shp::fail(__FILE__, __LINE__, __func__, (msg))
#define SHP_ASSERT(cond, msg) do { if (!(cond)) SHP_FAIL(msg); } while (0)

namespace shp {

[[noreturn]] void fail(const char *file, int line, const char *func,
                       const char *msg);

constexpr int MB = 8;        // m values per phase block
constexpr int kTabPad = 8;   // zero entries appended to every table row
constexpr int kAlmSlackRows = 16;   // rows after the last canonical a_lm row

struct RingDev {
  int64_t ofs;     // index of the first pixel in the map
  int nph;         // pixels in the ring (<=0: absent)
  int stride;      // index step between pixels
  double phi0;     // azimuth of the first pixel
  double weight;   // quadrature weight (USE_WEIGHTS)
};

__host__ __device__ inline int64_t tri_ofs(int m, int lmax)
{
  return (int64_t)m * (lmax + 1) - ((int64_t)m * (m - 1)) / 2;
}
__host__ __device__ inline int64_t tab_ofs(int m, int lmax)
{
  return tri_ofs(m, lmax) + (int64_t)m * kTabPad;
}
__host__ __device__ inline int64_t phase_index(int comp, int m, int slot,
                                               int hemi, int nslots, int nmb)
{
  return ((((int64_t)comp * nmb + m / MB) * nslots + slot) * 2 + hemi) * MB +
         (m % MB);
}

// exponent scaling (libsharp/sharp_ylmgen_c.h:39-42): a value v with scale s
// represents v * 2^(800*s)
constexpr double kFBig = 0x1p+800, kFSmall = 0x1p-800;
constexpr double kFTol = 0x1p-60, kFBigHalf = 0x1p+400;

} // namespace shp

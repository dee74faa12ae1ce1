// support_abi.cpp -- the rest of libsharp's exported interface that its own
// programs call: the start-up banner (libsharp/sharp_announce.c:45-98) and
// the host-side Y_lm generator (libsharp/sharp_ylmgen_c.c:44-232), on the
// layout of libsharp/sharp_ylmgen_c.h:47-73.  The generator's tables come
// from the same per-m routines as the engine's device tables
// (ylm_tables.hpp).  Memory the caller frees (get_norm, get_d1norm) comes from
// malloc, which libsharp's DEALLOC/free accept.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "sharp_announce.h"
#include "sharp_ylmgen_c.h"
#include "ylm_tables.hpp"

namespace shp {

template <typename T> static T *galloc(size_t n)
{
  T *p = static_cast<T *>(std::malloc(std::max<size_t>(n, 1) * sizeof(T)));
  if (!p) SHP_FAIL("malloc() failed");
  return p;
}

} // namespace shp

using namespace shp;

extern "C" {

void sharp_announce(const char *name)
{
  const size_t n = std::strlen(name);
  std::string bar(n + 2, '-');
  std::printf("\n+%s+\n| %s |\n+%s+\n\n", bar.c_str(), name, bar.c_str());
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
  std::printf("Engine: HIP spherical harmonic transforms for gfx950 (MI355X)\n");
  if (ndev == 0) std::printf("GPUs: none visible\n");
  for (int d = 0; d < ndev; ++d) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, d) == hipSuccess)
      std::printf("GPU %d: %s (%s), %d compute units, %.0f GB\n", d, p.name, p.gcnArchName,
                  p.multiProcessorCount, p.totalGlobalMem / 1e9);
  }
  std::printf("\n");
}

void sharp_module_startup(const char *name, int argc, int argc_expected,
                          const char *argv_expected, int verbose)
{
  if (verbose) sharp_announce(name);
  if (argc == argc_expected) return;
  if (verbose) std::fprintf(stderr, "Usage: %s %s\n", name, argv_expected);
  std::exit(1);
}

void sharp_Ylmgen_init(sharp_Ylmgen_C *gen, int l_max, int m_max, int spin)
{
  std::memset(gen, 0, sizeof(*gen));
  gen->lmax = l_max;
  gen->mmax = m_max;
  if (spin < 0) SHP_FAIL("incorrect spin: must be nonnegative");
  if (l_max < spin) SHP_FAIL("incorrect l_max: must be >= spin");
  if (l_max < m_max) SHP_FAIL("incorrect l_max: must be >= m_max");
  gen->s = spin;
  const int ncf = sharp_maxscale - sharp_minscale + 1;
  gen->cf = galloc<double>(ncf);
  for (int i = 0; i < ncf; ++i) {
    // 2^(800 (i + minscale)) by repeated scaling from 1
    double v = 1.;
    for (int k = 0; k < i + sharp_minscale; ++k) v *= sharp_fbig;
    for (int k = 0; k > i + sharp_minscale; --k) v *= sharp_fsmall;
    gen->cf[i] = v;
  }
  gen->m = -1;
  if (spin == 0) {
    gen->rf = galloc<sharp_ylmgen_dbl2>(l_max + 1);
    gen->mfac = galloc<double>(m_max + 1);
    ylm_mfac(m_max, gen->mfac);
    std::vector<double> root, iroot;
    ylm_roots(l_max, root, iroot);
    gen->root = galloc<double>(root.size());
    gen->iroot = galloc<double>(iroot.size());
    std::memcpy(gen->root, root.data(), root.size() * sizeof(double));
    std::memcpy(gen->iroot, iroot.data(), iroot.size() * sizeof(double));
    return;
  }
  gen->m = gen->mlo = gen->mhi = -1234567890;
  gen->fx = galloc<sharp_ylmgen_dbl3>(l_max + 2);
  std::memset(gen->fx, 0, (l_max + 2) * sizeof(sharp_ylmgen_dbl3));
  const SpinFactors f(l_max);
  gen->inv = galloc<double>(f.inv.size());
  gen->flm1 = galloc<double>(f.flm1.size());
  gen->flm2 = galloc<double>(f.flm2.size());
  std::memcpy(gen->inv, f.inv.data(), f.inv.size() * sizeof(double));
  std::memcpy(gen->flm1, f.flm1.data(), f.flm1.size() * sizeof(double));
  std::memcpy(gen->flm2, f.flm2.data(), f.flm2.size() * sizeof(double));
  gen->prefac = galloc<double>(m_max + 1);
  gen->fscale = galloc<int>(m_max + 1);
  ylm_spin_prefactors(l_max, m_max, spin, gen->prefac, gen->fscale);
}

void sharp_Ylmgen_destroy(sharp_Ylmgen_C *gen)
{
  for (void *p : {(void *)gen->cf, (void *)gen->rf, (void *)gen->mfac, (void *)gen->root,
                  (void *)gen->iroot, (void *)gen->fx, (void *)gen->prefac,
                  (void *)gen->fscale, (void *)gen->flm1, (void *)gen->flm2,
                  (void *)gen->inv})
    std::free(p);
  gen->cf = nullptr;
  gen->rf = nullptr; gen->mfac = nullptr; gen->root = nullptr; gen->iroot = nullptr;
  gen->fx = nullptr; gen->prefac = nullptr; gen->fscale = nullptr;
  gen->flm1 = nullptr; gen->flm2 = nullptr; gen->inv = nullptr;
}

void sharp_Ylmgen_prepare(sharp_Ylmgen_C *gen, int m)
{
  if (m == gen->m) return;
  if (m < 0) SHP_FAIL("incorrect m");
  gen->m = m;
  if (gen->s == 0) {
    ylm_spin0_coeffs(m, gen->lmax, gen->root, gen->iroot, &gen->rf[m].f[0]);
    return;
  }
  const int s = gen->s, mlo = std::min(m, s), mhi = std::max(m, s);
  // the fx entries depend on (m, s) only
  if (gen->mlo != mlo || gen->mhi != mhi)
    ylm_spin_coeffs(m, s, gen->lmax, gen->flm1, gen->flm2, gen->inv, &gen->fx[0].f[0], 0);
  gen->mlo = mlo;
  gen->mhi = mhi;
  // powers of cos(theta/2) and sin(theta/2) in the start values, and their
  // signs (libsharp/sharp_ylmgen_c.c:196-207)
  gen->preMinus_p = gen->preMinus_m = 0;
  if (mhi == m) {
    gen->cosPow = mhi + s;
    gen->sinPow = mhi - s;
    gen->preMinus_p = gen->preMinus_m = (mhi - s) & 1;
  } else {
    gen->cosPow = mhi + m;
    gen->sinPow = mhi - m;
    gen->preMinus_m = (mhi + m) & 1;
  }
}

double *sharp_Ylmgen_get_norm(int lmax, int spin)
{
  double *n = galloc<double>(lmax + 1);
  ylm_norm(lmax, spin, false, n);
  return n;
}

double *sharp_Ylmgen_get_d1norm(int lmax)
{
  double *n = galloc<double>(lmax + 1);
  ylm_norm(lmax, 1, true, n);
  return n;
}

} // extern "C"

// noalloc.hip -- the fixed work memory of sharp_hip_execute_device_noalloc
// (an entry point with no counterpart in the reference, for callers that
// cannot allocate device memory).
//
// `scratch` is a file-scope device array: every GPU that loads this code
// object holds its own copy, at its own address, until the process ends.  In
// host code the name `scratch` is only a stand-in for the symbol; the device
// address of the current GPU's copy comes from hipGetSymbolAddress, done once
// per device by noalloc_setup() (called from sharp_hip_prepare), and only
// that address is ever handed to kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>

#include "kernels.hpp"

namespace shp {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

// device -> address of its copy of `scratch` (guarded by the engine's mutex)
static std::map<int, void *> g_scratch_ptr;

void noalloc_setup()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) SHP_FAIL("hipGetDevice failed");
  if (g_scratch_ptr.count(dev)) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    char b[200];
    std::snprintf(b, sizeof(b), "hipGetSymbolAddress of the scratch area failed: %s",
                  hipGetErrorString(e));
    SHP_FAIL(b);
  }
  g_scratch_ptr[dev] = p;
}

void *noalloc_scratch(int device)
{
  auto it = g_scratch_ptr.find(device);
  return it == g_scratch_ptr.end() ? nullptr : it->second;
}

// One lane walks the m values (a few thousand at most): a prefix sum too
// small to be worth spreading.
__global__ void k_bs_base(int64_t *base, int lmax, int mmax, int spin, int ncol)
{
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t t = 0;
  for (int m = 0; m <= mmax; ++m) {
    base[m] = t;
    t += (int64_t)3 * ncol * bs_pairs(lmax, m > spin ? m : spin);
  }
}

void launch_bs_base(int64_t *base, int lmax, int mmax, int spin, int ncol, hipStream_t st)
{
  hipLaunchKernelGGL(k_bs_base, dim3(1), dim3(64), 0, st, base, lmax, mmax, spin, ncol);
}

} // namespace shp

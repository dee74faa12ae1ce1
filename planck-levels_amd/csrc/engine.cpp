// engine.cpp -- host side of the MI355X SHT engine: job normalisation and
// validation, per-device caches of geometry / recurrence tables, host<->device
// staging for the libsharp ABI, and the sequence of kernel launches.
//
// Job normalisation and checks follow libsharp/sharp.c:238-251,762-880 and
// libsharp/sharp_ylmgen_c.c:44-52; the recurrence tables are built exactly as
// libsharp/sharp_ylmgen_c.c:44-232 builds them (same operations in the same
// order), so both implementations start from the same coefficients (up to
// the regrouping of products -ffast-math allows the reference build); the
// spin-s f1 is stored premultiplied by f0 (one product more, see rec_step).
//
// Pipeline per call (one HIP stream, no host synchronisation inside):
//   synthesis: alm_gather -> Legendre synthesis (per transform group)
//              -> phase2ring (writes the map, adds with SHARP_ADD)
//   analysis:  ring2phase -> Legendre analysis (per transform group)
//              -> alm_scatter (writes the a_lm, adds with SHARP_ADD)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "sharp_lowlevel.h"
#include "sharp_hip.h"
#include "shp_internal.hpp"
#include "kernels.hpp"
#include "ylm_tables.hpp"

namespace shp {

#define HIP_CHECK(expr)                                                        \
  do {                                                                         \
    const hipError_t e_ = (expr);                                              \
    if (e_ != hipSuccess) {                                                    \
      char b_[320];                                                            \
      std::snprintf(b_, sizeof(b_), "HIP error '%s' in %s",                    \
                    hipGetErrorString(e_), #expr);                             \
      SHP_FAIL(b_);                                                            \
    }                                                                          \
  } while (0)


// ---------------------------------------------------------------------------
// device memory helpers
// ---------------------------------------------------------------------------
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  bool pooled = false;   // from hipMallocAsync
  // A buffer of at least `bytes`.  Buffers that belong to one launch stream
  // (ordered) grow in stream order: the old one is freed on that stream after
  // the work already queued there, with no device-wide wait.  Others wait
  // for the device before the old buffer is freed.
  void *get(size_t bytes, hipStream_t st = nullptr, bool ordered = false)
  {
    if (bytes == 0) bytes = 16;
    if (bytes > cap) {
      if (p) {
        if (ordered) {
          HIP_CHECK(hipFreeAsync(p, st));
        } else {
          HIP_CHECK(hipDeviceSynchronize());
          HIP_CHECK(pooled ? hipFreeAsync(p, nullptr) : hipFree(p));
          if (pooled) HIP_CHECK(hipStreamSynchronize(nullptr));
        }
      }
      p = nullptr;
      if (ordered) HIP_CHECK(hipMallocAsync(&p, bytes, st));
      else HIP_CHECK(hipMalloc(&p, bytes));
      pooled = ordered;
      cap = bytes;
    }
    return p;
  }
  void release()
  {
    if (p) {
      if (pooled) {
        (void)hipFreeAsync(p, nullptr);
        (void)hipStreamSynchronize(nullptr);
      } else {
        (void)hipFree(p);
      }
    }
    p = nullptr;
    cap = 0;
    pooled = false;
  }
};

// ---------------------------------------------------------------------------
// host <-> device staging of the host-pointer sharp_execute
// ---------------------------------------------------------------------------
// Copies between the caller's pageable arrays and pinned memory with several
// host threads (one thread moves ~28 GB/s, eight ~100 GB/s on the MI355X
// hosts, where a pageable hipMemcpy reaches 28 GB/s to and 17 GB/s from the
// device and a pinned one 57 GB/s each way).  The workers live for the
// process; they are detached and never joined.
class CopyPool {
public:
  explicit CopyPool(int nthreads) : n_(std::max(1, nthreads))
  {
    for (int i = 1; i < n_; ++i) std::thread([this, i] { work(i); }).detach();
  }
  // dst[0..bytes) = src[0..bytes), split over the workers and the caller
  void copy(void *dst, const void *src, size_t bytes)
  {
    if (n_ == 1 || bytes < (size_t(1) << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      dst_ = static_cast<char *>(dst);
      src_ = static_cast<const char *>(src);
      bytes_ = bytes;
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    part(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

private:
  void part(int i)
  {
    // stripes of whole 4 KiB pages
    const size_t pages = (bytes_ + 4095) / 4096;
    const size_t a = std::min(bytes_, pages * i / n_ * 4096);
    const size_t b = std::min(bytes_, pages * (i + 1) / n_ * 4096);
    if (b > a) std::memcpy(dst_ + a, src_ + a, b - a);
  }
  void work(int i)
  {
    unsigned long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      part(i);
      std::lock_guard<std::mutex> lk(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  const int n_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t bytes_ = 0;
  int pending_ = 0;
  unsigned long gen_ = 0;
};

static CopyPool &copy_pool()
{
  static CopyPool *pool = [] {
    int n = 8;
    if (const char *e = std::getenv("SHARP_HIP_STAGE_THREADS")) n = std::atoi(e);
    const int hw = (int)std::thread::hardware_concurrency();
    if (hw > 0) n = std::min(n, hw);
    return new CopyPool(std::max(1, std::min(n, 64)));
  }();
  return *pool;
}

// A ring of pinned chunks between the caller's arrays and HBM: the host
// threads fill (or drain) one chunk while the DMA engine moves another, on a
// copy stream ordered against the launch stream with events.
struct HostStage {
  static constexpr int kSlots = 4;
  static constexpr size_t kChunk = size_t(64) << 20;
  char *pinned = nullptr;
  hipEvent_t ev[kSlots] = {}, mark = nullptr, done[2] = {};
  bool busy[kSlots] = {};
  hipStream_t copy = nullptr;
  int next = 0;
  void init()
  {
    if (pinned) return;
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&pinned), kSlots * kChunk));
    for (auto &e : ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&mark, hipEventDisableTiming));
    for (auto &e : done) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CHECK(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
  }
  void release()
  {
    if (!pinned) return;
    (void)hipStreamSynchronize(copy);
    (void)hipHostFree(pinned);
    for (auto &e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(mark);
    for (auto &e : done) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(copy);
    pinned = nullptr;
    for (auto &b : busy) b = false;
  }
  struct Seg { void *dst; const void *src; size_t bytes; };
  // host -> device; work enqueued on `st` afterwards sees the data
  void upload(const std::vector<Seg> &segs, hipStream_t st)
  {
    init();
    for (const Seg &g : segs)
      for (size_t off = 0; off < g.bytes; off += kChunk) {
        const size_t n = std::min(kChunk, g.bytes - off);
        const int k = next;
        next = (next + 1) % kSlots;
        if (busy[k]) HIP_CHECK(hipEventSynchronize(ev[k]));
        char *slot = pinned + k * kChunk;
        copy_pool().copy(slot, static_cast<const char *>(g.src) + off, n);
        HIP_CHECK(hipMemcpyAsync(static_cast<char *>(g.dst) + off, slot, n,
                                 hipMemcpyHostToDevice, copy));
        HIP_CHECK(hipEventRecord(ev[k], copy));
        busy[k] = true;
      }
    HIP_CHECK(hipEventRecord(mark, copy));
    HIP_CHECK(hipStreamWaitEvent(st, mark, 0));
  }
  // device -> host after the work enqueued on `st`; returns when done
  void download(const std::vector<Seg> &segs, hipStream_t st)
  {
    init();
    HIP_CHECK(hipEventRecord(mark, st));
    download_after(segs, mark);
  }
  // device -> host once `ready` has completed; returns when done
  void download_after(const std::vector<Seg> &segs, hipEvent_t ready)
  {
    init();
    HIP_CHECK(hipStreamWaitEvent(copy, ready, 0));
    struct Chunk { char *dst; const char *src; size_t n; int slot; };
    std::vector<Chunk> ch;
    for (const Seg &g : segs)
      for (size_t off = 0; off < g.bytes; off += kChunk)
        ch.push_back(Chunk{static_cast<char *>(g.dst) + off,
                           static_cast<const char *>(g.src) + off,
                           std::min(kChunk, g.bytes - off), -1});
    auto issue = [&](size_t i) {
      const int k = next;
      next = (next + 1) % kSlots;
      if (busy[k]) HIP_CHECK(hipEventSynchronize(ev[k]));
      ch[i].slot = k;
      HIP_CHECK(hipMemcpyAsync(pinned + k * kChunk, ch[i].src, ch[i].n,
                               hipMemcpyDeviceToHost, copy));
      HIP_CHECK(hipEventRecord(ev[k], copy));
      busy[k] = true;
    };
    const size_t ahead = kSlots - 1;
    for (size_t i = 0; i < ch.size() && i < ahead; ++i) issue(i);
    for (size_t i = 0; i < ch.size(); ++i) {
      const int k = ch[i].slot;
      HIP_CHECK(hipEventSynchronize(ev[k]));
      busy[k] = false;
      copy_pool().copy(ch[i].dst, pinned + k * kChunk, ch[i].n);
      if (i + ahead < ch.size()) issue(i + ahead);
    }
  }
};

template <typename T> static T *upload(const std::vector<T> &v)
{
  T *d = nullptr;
  HIP_CHECK(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (!v.empty())
    HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// ---------------------------------------------------------------------------
// cached device objects
// ---------------------------------------------------------------------------
struct GeomPlan {
  std::vector<char> key;     // the ring data this plan was built from
  int nslots = 0, nphmax = 0;
  std::vector<int> order;    // slot -> index into geom_info->pair
  std::vector<double> hcth, hsth;
  double *cth = nullptr, *sth = nullptr;
  int *ispair = nullptr;
  RingDev *ring = nullptr;
  struct Mlim {
    int lmax, spin;
    std::vector<int> h;
    int *d = nullptr;
    // largest mlim of each group of mix_group_slots(spin) slots (mixed analysis)
    std::vector<int> hgmax;
    int *gmax = nullptr;
  };
  // held by pointer: JobStates keep references to entries while later jobs
  // on the same geometry add more
  std::vector<std::unique_ptr<Mlim>> mlims;
  // ring transforms: FFT plans grouped by LDS class, the rings too long for
  // any class (direct summation), and the list of all rings (SHARP_NO_FFT)
  RingFftPlan *fft[kFftKinds * kNumFftClasses] = {};   // [kFftKinds*class + kind]
  std::vector<RingFftPlan> hfft[kFftKinds * kNumFftClasses];
  int nfft[kFftKinds * kNumFftClasses] = {};
  double2 *btab = nullptr;   // Bluestein chirp spectra
  // rings too long for every class: global-memory FFTs (long_plans), with
  // their Bluestein kernel spectra (ltab); lg_max: log2 of the largest length
  LongRingPlan *longp = nullptr;
  std::vector<LongRingPlan> hlong;
  int nlong = 0, lg_max = 0;
  double2 *ltab = nullptr;
  int64_t npix_long = 0;
  int *dft_list = nullptr, ndft = 0;
  int *all_list = nullptr, nall = 0;
  int64_t npix_fft = 0, npix_all = 0;   // pixels in FFT rings / in all rings
  ~GeomPlan()
  {
    (void)hipFree(cth); (void)hipFree(sth); (void)hipFree(ispair); (void)hipFree(ring);
    for (auto &m : mlims) { (void)hipFree(m->d); (void)hipFree(m->gmax); }
    for (auto *p : fft) (void)hipFree(p);
    (void)hipFree(btab);
    (void)hipFree(longp);
    (void)hipFree(ltab);
    (void)hipFree(dft_list);
    (void)hipFree(all_list);
  }
};

struct YlmPlan {
  int lmax, mmax, spin;
  double *tab = nullptr;     // rows at tab_ofs(m, lmax)
  double *mfac = nullptr;    // spin 0
  double *prefac = nullptr;  // spin s
  int *fscale = nullptr;
  // built for the first mixed-precision job: the table of
  // lambda_l / s_l (all b_l or f2_l = 1) and s_{m,l} at tri_ofs(m, lmax) + l - m
  double *tab1 = nullptr, *lsc = nullptr;
  ~YlmPlan()
  {
    (void)hipFree(tab); (void)hipFree(mfac); (void)hipFree(prefac); (void)hipFree(fscale);
    (void)hipFree(tab1); (void)hipFree(lsc);
  }
};

struct NormPlan {
  int lmax, spin, deriv;
  double *d = nullptr;
  ~NormPlan() { (void)hipFree(d); }
};

// word offsets of the BF16 pieces of each m (bs_pairs)
struct BsPlan {
  int lmax, mmax, spin, ncol;
  int64_t total = 0;
  int64_t *d = nullptr;
  ~BsPlan() { (void)hipFree(d); }
};

struct AlmPlan {
  std::vector<char> key;
  int64_t *mvstart_by_m = nullptr;
  ~AlmPlan() { (void)hipFree(mvstart_by_m); }
};

// Opt-in per-stage timing: HIP events recorded on the launch stream around
// each stage, with the algorithmic work of that launch (libsharp's flop
// model for the Legendre stage, compulsory HBM bytes for the others).
struct Profiler {
  bool on = false;
  struct Pending {
    int stage;
    hipEvent_t a, b;
    double flops, bytes;
  };
  struct Stat {
    std::string name;
    double ms = 0., flops = 0., bytes = 0.;
    long long launches = 0;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::vector<Stat> stats;
  hipEvent_t event()
  {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  int stage(const char *name)
  {
    for (size_t i = 0; i < stats.size(); ++i)
      if (stats[i].name == name) return (int)i;
    stats.push_back(Stat{name});
    return (int)stats.size() - 1;
  }
  void drain()
  {
    for (auto &p : pending) {
      HIP_CHECK(hipEventSynchronize(p.b));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
      Stat &st = stats[p.stage];
      st.ms += ms;
      st.flops += p.flops;
      st.bytes += p.bytes;
      st.launches += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
};

struct Ctx {
  int device = 0;
  Profiler prof;
  int max_lds = 65536;       // bytes of LDS one workgroup may allocate
  hipStream_t stream = nullptr;
  // cached plans, shared with the jobs of live batches: sharp_hip_release
  // drops the cache's references, a batch keeps its own until destroyed
  std::vector<std::shared_ptr<GeomPlan>> geoms;
  std::vector<std::shared_ptr<YlmPlan>> ylms;
  std::vector<std::shared_ptr<NormPlan>> norms;
  std::vector<std::shared_ptr<AlmPlan>> alms;
  std::vector<std::shared_ptr<BsPlan>> bsplans;
  // work buffers per launch stream: calls on different streams do not share
  // them and may run concurrently
  struct StreamBufs {
    DevBuf phase, almtmp, part, resid, lfac, bsplit, lscratch;
    // the ring FFT stage's two extra streams (created on first use)
    FftFork fork{};
    bool has_fork = false;
    ~StreamBufs()
    {
      if (!has_fork) return;
      for (int i = 0; i < 2; ++i) {
        (void)hipStreamDestroy(fork.aux[i]);
        (void)hipEventDestroy(fork.join[i]);
      }
      (void)hipEventDestroy(fork.fork);
    }
  };
  std::map<hipStream_t, std::unique_ptr<StreamBufs>> sbufs;
  StreamBufs &bufs(hipStream_t st)
  {
    auto &p = sbufs[st];
    if (!p) p = std::make_unique<StreamBufs>();
    return *p;
  }
  DevBuf stage_alm, stage_map;
  HostStage host;   // pinned chunks of the host-pointer sharp_execute
  // pinned staging of the per-l factors; lfac_ev marks the end of the last
  // copy out of it
  double *lfac_host = nullptr;
  size_t lfac_cap = 0;
  hipEvent_t lfac_ev = nullptr;
};

static std::mutex g_mutex;
static std::map<int, std::unique_ptr<Ctx>> g_ctx;
// Set (under g_mutex) while sharp_hip_execute_device_noalloc runs: a table
// that is not cached yet is then an error instead of being built.
static bool g_noalloc = false;

// long-ring FFT slots of sharp_hip_execute_device_noalloc
constexpr int kNoallocLongSlots = 8;

static void need_build(const char *what)
{
  if (!g_noalloc) return;
  char b[256];
  std::snprintf(b, sizeof(b),
                "sharp_hip_execute_device_noalloc: the %s of this job is not prepared "
                "(call sharp_hip_prepare with this geometry, a_lm layout and spin first)",
                what);
  SHP_FAIL(b);
}

static Ctx &context()
{
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    SHP_FAIL("no GPU device available: the HIP engine cannot run here");
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  auto it = g_ctx.find(dev);
  if (it != g_ctx.end()) return *it->second;
  auto c = std::make_unique<Ctx>();
  c->device = dev;
  HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
          hipSuccess && lds > 0)
    c->max_lds = lds;
  Ctx &ref = *c;
  g_ctx[dev] = std::move(c);
  return ref;
}

// --- ring FFT plans -----------------------------------------------------------
// Radices for a length: powers of 8, then one 4 or 2, then 3, 5, 7.  Returns
// the cofactor left over (1 when the length is 7-smooth).
static int radices(int N, std::vector<int> &rad)
{
  rad.clear();
  int n = N;
  while (n % 8 == 0) { rad.push_back(8); n /= 8; }
  while (n % 4 == 0) { rad.push_back(4); n /= 4; }
  while (n % 2 == 0) { rad.push_back(2); n /= 2; }
  for (int p : {3, 5, 7})
    while (n % p == 0) { rad.push_back(p); n /= p; }
  return n;
}

// Plan for a ring of n pixels; returns its LDS class (buffer of 512<<class
// complex doubles) or -1 when no class is large enough.
static int fft_plan(int rid, int n, RingFftPlan &pl)
{
  std::memset(&pl, 0, sizeof(pl));
  pl.ring = rid;
  pl.n = n;
  pl.N = (n & 1) ? n : n / 2;
  std::vector<int> rad;
  size_t need;
  if (radices(pl.N, rad) == 1) {
    pl.bluestein = 0;
    need = (size_t)pl.N;
  } else {
    // convolution length: the smallest power of two >= 2N-1 and >= 64, so
    // that every pass is of radix 8, 4 or 2 and the first and the last are
    // of radix 8 (measured faster overall than the smallest 7-smooth length,
    // whose radix 3, 5 and 7 passes cost more per point)
    pl.bluestein = 1;
    size_t M = 64;
    while (M < 2 * (size_t)pl.N - 1) M <<= 1;
    radices((int)M, rad);
    rad.erase(rad.begin());
    rad.push_back(8);
    // M = 8192 (the class of 16 values per thread): 8, 16, 8, 8 (four LDS
    // passes instead of five)
    if (M == 8192) rad = {8, 16, 8, 8};
    pl.M = (int)M;
    need = M;
  }
  if (rad.size() > sizeof(pl.rad)) return -1;
  pl.nrad = (int)rad.size();
  for (size_t i = 0; i < rad.size(); ++i) pl.rad[i] = (int8_t)rad[i];
  for (int cl = 0; cl < kNumFftClasses; ++cl)
    if (need <= (size_t)(512 << cl)) return cl;
  return -1;
}

// Plan of a ring too long for every LDS class: one FFT of length 2^lg,
// n itself or (Bluestein) the smallest power of two >= 2n-1.  False for
// rings longer than 2^21 pixels (the direct sums take those).
static bool long_plan(int rid, int n, LongRingPlan &pl)
{
  pl = LongRingPlan{rid, n, 0, 0, -1};
  if (n < 2) return false;
  const bool pow2 = (n & (n - 1)) == 0;
  const int64_t target = pow2 ? n : 2 * (int64_t)n - 1;
  while (((int64_t)1 << pl.lg) < target) ++pl.lg;
  pl.bluestein = pow2 ? 0 : 1;
  return pl.lg <= 22;
}

// the LDS class of a ring of n pixels on a device with max_lds bytes of LDS
// per workgroup (-1: none fits)
static int ring_class(int rid, int n, int max_lds, RingFftPlan &pl)
{
  int cl = fft_plan(rid, n, pl);
  // the class's FFT buffer (with its pad of one entry in 8) plus room to
  // stage 512 phases (ring_fft.hip)
  if (cl >= 0 && (size_t)((512 << cl) + (64 << cl) + 512) * sizeof(double2) + 256 > (size_t)max_lds)
    cl = -1;
  return cl;
}

// --- geometry ---------------------------------------------------------------
static std::vector<char> geom_key(const sharp_geom_info *g)
{
  std::vector<char> k;
  k.reserve(size_t(g->npairs) * 2 * sizeof(sharp_ringinfo) + 8);
  auto put = [&](const void *p, size_t n) {
    const char *c = static_cast<const char *>(p);
    k.insert(k.end(), c, c + n);
  };
  put(&g->npairs, sizeof(int));
  for (int i = 0; i < g->npairs; ++i) {
    put(&g->pair[i].r1, sizeof(sharp_ringinfo));
    if (g->pair[i].r2.nph > 0)
      put(&g->pair[i].r2, sizeof(sharp_ringinfo));
    else {
      const int none = -1;
      put(&none, sizeof(int));
    }
  }
  return k;
}

static GeomPlan &get_geom(Ctx &c, const sharp_geom_info *g)
{
  std::vector<char> key = geom_key(g);
  for (auto &p : c.geoms)
    if (p->key == key) return *p;
  need_build("geometry");
  auto gp = std::make_shared<GeomPlan>();
  gp->key = std::move(key);
  const int P = g->npairs;
  gp->nslots = P;
  gp->order.resize(P);
  std::iota(gp->order.begin(), gp->order.end(), 0);
  // device order: by sin(theta) of the northern ring, so that wavefronts see
  // rings with similar m cutoffs and similar exponent-scaling behaviour
  std::stable_sort(gp->order.begin(), gp->order.end(), [g](int a, int b) {
    return g->pair[a].r1.sth < g->pair[b].r1.sth;
  });
  std::vector<int> ispair(P);
  std::vector<RingDev> ring(2 * size_t(P));
  gp->hcth.resize(P);
  gp->hsth.resize(P);
  int nphmax = 0;
  for (int k = 0; k < P; ++k) {
    const sharp_ringpair &pr = g->pair[gp->order[k]];
    gp->hcth[k] = pr.r1.cth;
    gp->hsth[k] = pr.r1.sth;
    ispair[k] = pr.r2.nph > 0;
    ring[2 * k] = RingDev{(int64_t)pr.r1.ofs, pr.r1.nph, pr.r1.stride, pr.r1.phi0,
                          pr.r1.weight};
    if (pr.r2.nph > 0)
      ring[2 * k + 1] = RingDev{(int64_t)pr.r2.ofs, pr.r2.nph, pr.r2.stride,
                                pr.r2.phi0, pr.r2.weight};
    else
      ring[2 * k + 1] = RingDev{0, -1, 0, 0., 0.};
    nphmax = std::max(nphmax, std::max(pr.r1.nph, pr.r2.nph));
  }
  gp->nphmax = nphmax;
  // ring FFT plans by LDS class; rings beyond every class use direct sums
  std::vector<RingFftPlan> cls[kFftKinds * kNumFftClasses];
  std::vector<int> dft, all;
  for (int rid = 0; rid < 2 * P; ++rid) {
    const int n = ring[rid].nph;
    if (n <= 0) continue;
    all.push_back(rid);
    RingFftPlan pl;
    const int cl = ring_class(rid, n, c.max_lds, pl);
    gp->npix_all += n;
    LongRingPlan lp;
    if (cl < 0 && long_plan(rid, n, lp)) {
      gp->hlong.push_back(lp);
      gp->npix_long += n;
    } else if (cl < 0) {
      dft.push_back(rid);
    } else {
      // radix set: 2 only 8, 1 only 8, 4, 2 (and 16), 0 any
      int rs = 2;
      for (int q = 0; q < pl.nrad; ++q) {
        if (pl.rad[q] == 4 || pl.rad[q] == 2 || pl.rad[q] == 16) rs = std::min(rs, 1);
        else if (pl.rad[q] != 8) rs = 0;
      }
      cls[kFftKinds * cl + pl.bluestein + 2 * rs].push_back(pl);
      gp->npix_fft += n;
    }
  }
  // Bluestein chirp spectra: one table entry per distinct (N, M)
  std::map<std::pair<int, int>, int64_t> boffs;
  int64_t btot = 0;
  FftClasses uniq{};
  // (the table kernels take every Bluestein plan of a class from slot
  // kFftKinds*class + 1, whatever its radices)
  std::vector<RingFftPlan> ucls[kFftKinds * kNumFftClasses];
  for (int c2 = 0; c2 < kFftKinds * kNumFftClasses; ++c2) {
    for (auto &pl : cls[c2]) {
      pl.boff = -1;
      if (!pl.bluestein) continue;
      auto key = std::make_pair(pl.N, pl.M);
      auto it = boffs.find(key);
      if (it == boffs.end()) {
        it = boffs.emplace(key, btot).first;
        // kernel spectrum, then the chirp: the halves 0..M/2 and 0..N/2 of
        // the two symmetric sequences
        btot += pl.M / 2 + 1 + pl.N / 2 + 1;
        RingFftPlan u = pl;
        u.boff = it->second;
        ucls[(c2 / kFftKinds) * kFftKinds + 1].push_back(u);
      }
      pl.boff = it->second;
    }
    // longest rings first: they take longest, and the rings that need the
    // phase staging of the synthesis fold (n <= mmax) come last
    std::stable_sort(cls[c2].begin(), cls[c2].end(),
                     [](const RingFftPlan &x, const RingFftPlan &y) { return x.n > y.n; });
    gp->nfft[c2] = (int)cls[c2].size();
    if (!cls[c2].empty()) gp->fft[c2] = upload(cls[c2]);
    gp->hfft[c2] = cls[c2];
  }
  if (gp->npix_fft > 0) ensure_twiddles(c.stream);
  if (btot > 0) {
    HIP_CHECK(hipMalloc(&gp->btab, (size_t)btot * sizeof(double2)));
    std::vector<RingFftPlan *> dev(kFftKinds * kNumFftClasses, nullptr);
    for (int c2 = 0; c2 < kFftKinds * kNumFftClasses; ++c2) {
      uniq.count[c2] = (int)ucls[c2].size();
      if (!ucls[c2].empty()) dev[c2] = upload(ucls[c2]);
      uniq.plans[c2] = dev[c2];
      uniq.hplans[c2] = ucls[c2].data();
    }
    launch_bluestein_table(uniq, gp->btab, c.stream);
    HIP_CHECK(hipStreamSynchronize(c.stream));
    for (auto *d : dev) (void)hipFree(d);
  }
  if (!gp->hlong.empty()) {
    // Bluestein kernel spectra of the long rings, one block per length
    std::map<int, int64_t> loffs;
    std::vector<LongRingPlan> uniq;
    int64_t ltot = 0;
    for (auto &lp : gp->hlong) {
      gp->lg_max = std::max(gp->lg_max, lp.lg);
      if (!lp.bluestein) continue;
      auto it = loffs.find(lp.n);
      if (it == loffs.end()) {
        it = loffs.emplace(lp.n, ltot).first;
        ltot += (int64_t)1 << lp.lg;
        LongRingPlan u = lp;
        u.boff = it->second;
        uniq.push_back(u);
      }
      lp.boff = it->second;
    }
    // longest first: they take longest
    std::stable_sort(gp->hlong.begin(), gp->hlong.end(),
                     [](const LongRingPlan &x, const LongRingPlan &y) { return x.lg > y.lg; });
    gp->nlong = (int)gp->hlong.size();
    gp->longp = upload(gp->hlong);
    if (ltot > 0) {
      HIP_CHECK(hipMalloc(&gp->ltab, (size_t)ltot * sizeof(double2)));
      int lgu = 0;
      for (auto &u : uniq) lgu = std::max(lgu, u.lg);
      const int nwg = std::min<int>((int)uniq.size(), 256);
      const int64_t slot = 2 * ((int64_t)1 << lgu);
      double2 *tmp = nullptr;
      HIP_CHECK(hipMalloc(&tmp, (size_t)nwg * slot * sizeof(double2)));
      LongRingPlan *du = upload(uniq);
      launch_long_table(du, (int)uniq.size(), tmp, slot, nwg, gp->ltab, c.stream);
      HIP_CHECK(hipStreamSynchronize(c.stream));
      (void)hipFree(du);
      (void)hipFree(tmp);
    }
  }
  gp->ndft = (int)dft.size();
  gp->dft_list = upload(dft);
  gp->nall = (int)all.size();
  gp->all_list = upload(all);
  gp->cth = upload(gp->hcth);
  gp->sth = upload(gp->hsth);
  gp->ispair = upload(ispair);
  gp->ring = upload(ring);
  c.geoms.push_back(std::move(gp));
  return *c.geoms.back();
}

static GeomPlan::Mlim &get_mlim(GeomPlan &gp, int lmax, int spin)
{
  for (auto &m : gp.mlims)
    if (m->lmax == lmax && m->spin == spin) return *m;
  need_build("per-ring m limit table");
  auto ml = std::make_unique<GeomPlan::Mlim>();
  ml->lmax = lmax;
  ml->spin = spin;
  ml->h.resize(gp.nslots);
  for (int k = 0; k < gp.nslots; ++k)
    ml->h[k] = sharp_get_mlim(lmax, spin, gp.hsth[k], gp.hcth[k]);
  ml->d = upload(ml->h);
  const int gs = mix_group_slots(spin);
  const int ng = (gp.nslots + gs - 1) / gs;
  ml->hgmax.assign(ng, -1);
  for (int k = 0; k < gp.nslots; ++k) ml->hgmax[k / gs] = std::max(ml->hgmax[k / gs], ml->h[k]);
  ml->gmax = upload(ml->hgmax);
  gp.mlims.push_back(std::move(ml));
  return *gp.mlims.back();
}

// --- recurrence tables (libsharp/sharp_ylmgen_c.c:44-198) --------------------
// the spin-s recurrence coefficients (f0, f0 f1, f2) of entry l at
// 3 (tab_ofs(m, lmax) + l - m)
static std::vector<double> spin_table(int lmax, int mmax, int s)
{
  std::vector<double> tab(3 * (size_t)tab_ofs(mmax + 1, lmax), 0.);
  const SpinFactors f(lmax);
  for (int m = 0; m <= mmax; ++m) {
    double *row = tab.data() + 3 * tab_ofs(m, lmax);
    ylm_spin_coeffs(m, s, lmax, f.flm1.data(), f.flm2.data(), f.inv.data(), row, m);
    // stored as f0*f1: the kernels form (cth -+ f1) f0 as one FMA
    for (int l = std::max(m, s); l < lmax; ++l) {
      double *e = row + 3 * (l + 1 - m);
      e[1] = e[0] * e[1];
    }
  }
  return tab;
}

// the spin-0 recurrence coefficients (a_l, b_l) of entry l at
// 2 (tab_ofs(m, lmax) + l - m)
static std::vector<double> spin0_table(int lmax, int mmax)
{
  std::vector<double> tab(2 * (size_t)tab_ofs(mmax + 1, lmax), 0.), root, iroot;
  ylm_roots(lmax, root, iroot);
  for (int m = 0; m <= mmax; ++m)
    ylm_spin0_coeffs(m, lmax, root.data(), iroot.data(), tab.data() + 2 * tab_ofs(m, lmax));
  return tab;
}

// Spin 0: entry l gives lambda_{l+1} = a_l cth lambda_l - b_l lambda_{l-1};
// with s_m = s_{m+1} = 1 and s_{l+1} = b_l s_{l-1}, mu_l = lambda_l / s_l obeys
// mu_{l+1} = a_l (s_l / s_{l+1}) cth mu_l - mu_{l-1} (s within [0.1, 1.2]).
// The mixed-precision kernels run it.
static void unit_table0(const YlmPlan &yp, std::vector<double> &tab, std::vector<double> &lsc)
{
  const int lmax = yp.lmax, mmax = yp.mmax;
  tab = spin0_table(lmax, mmax);
  for (int m = 0; m <= mmax; ++m) {
    double *row = tab.data() + 2 * (tab_ofs(m, lmax) - m);
    double *sl = lsc.data() + tri_ofs(m, lmax) - m;
    for (int l = m + 2; l <= lmax; ++l) {
      const double b = row[2 * (l - 1) + 1];
      if (!(b > 0.)) SHP_FAIL("spin-0 recurrence coefficient b is not positive");
      sl[l] = b * sl[l - 2];
    }
    for (int l = m; l < lmax; ++l) row[2 * l] *= sl[l] / sl[l + 1];
    for (int l = m; l <= lmax; ++l) row[2 * l + 1] = 1.;
  }
}

// The recurrence lambda_l = c_l lambda_{l-1} - f2_l lambda_{l-2}, c_l =
// cth f0_l -+ g_l, rewritten for mu_l = lambda_l / s_l with s_l = f2_l s_{l-2}
// and s = 1 at l = mhi, mhi+1 (the f2 of mhi+1 multiplies lambda_{mhi-1} = 0):
// mu_l = c_l (s_{l-1} / s_l) mu_{l-1} - mu_{l-2}, one multiplication less per
// step.  The mixed-precision kernels run it; the a_lm gather and scatter apply
// s_{m,l}.  s lies in (0, 1]: it is a product of factors f2_l <= l/(l+1) of
// one parity ((l^2-m^2)/((l+1)^2-m^2) <= l^2/(l+1)^2, likewise for s); its
// minimum is 0.017 at lmax 4096 and 0.008 at lmax 16384.
static void ensure_unit_table(YlmPlan &yp)
{
  if (yp.tab1) return;
  need_build("mixed-precision recurrence table");
  const int lmax = yp.lmax, mmax = yp.mmax, s = yp.spin;
  std::vector<double> tab, lsc((size_t)tri_ofs(mmax + 1, lmax), 1.);
  if (s == 0) {
    unit_table0(yp, tab, lsc);
    yp.tab1 = upload(tab);
    yp.lsc = upload(lsc);
    return;
  }
  tab = spin_table(lmax, mmax, s);
  for (int m = 0; m <= mmax; ++m) {
    double *row = tab.data() + 3 * (tab_ofs(m, lmax) - m);
    double *sl = lsc.data() + tri_ofs(m, lmax) - m;
    const int mhi = std::max(m, s);
    for (int l = mhi + 2; l <= lmax; ++l) {
      const double f2 = row[3 * l + 2];
      if (!(f2 > 0.)) SHP_FAIL("spin recurrence coefficient f2 is not positive");
      sl[l] = f2 * sl[l - 2];
    }
    for (int l = mhi + 1; l <= lmax; ++l) {
      double *e = row + 3 * l;
      const double r = sl[l - 1] / sl[l];
      e[0] *= r;
      e[1] *= r;
      e[2] = 1.;
    }
  }
  yp.tab1 = upload(tab);
  yp.lsc = upload(lsc);
}

static YlmPlan &get_ylm(Ctx &c, int lmax, int mmax, int spin)
{
  for (auto &p : c.ylms)
    if (p->lmax == lmax && p->mmax == mmax && p->spin == spin) return *p;
  need_build("recurrence table");
  auto yp = std::make_shared<YlmPlan>();
  yp->lmax = lmax;
  yp->mmax = mmax;
  yp->spin = spin;
  if (spin == 0) {
    std::vector<double> tab = spin0_table(lmax, mmax);
    std::vector<double> mfac(mmax + 1);
    ylm_mfac(mmax, mfac.data());
    yp->tab = upload(tab);
    yp->mfac = upload(mfac);
  } else {
    std::vector<double> tab = spin_table(lmax, mmax, spin);
    std::vector<double> prefac(mmax + 1);
    std::vector<int> fscale(mmax + 1);
    ylm_spin_prefactors(lmax, mmax, spin, prefac.data(), fscale.data());
    yp->tab = upload(tab);
    yp->prefac = upload(prefac);
    yp->fscale = upload(fscale);
  }
  c.ylms.push_back(std::move(yp));
  return *c.ylms.back();
}

// norm_l (libsharp/sharp_ylmgen_c.c:200-232)
static std::vector<double> norm_values(int lmax, int spin, int deriv)
{
  std::vector<double> n(lmax + 1);
  ylm_norm(lmax, spin, deriv != 0, n.data());
  return n;
}

static double *get_norm(Ctx &c, int lmax, int spin, int deriv)
{
  for (auto &p : c.norms)
    if (p->lmax == lmax && p->spin == spin && p->deriv == deriv) return p->d;
  need_build("norm_l table");
  const std::vector<double> n = norm_values(lmax, spin, deriv);
  auto np = std::make_shared<NormPlan>();
  np->lmax = lmax;
  np->spin = spin;
  np->deriv = deriv;
  np->d = upload(n);
  c.norms.push_back(std::move(np));
  return c.norms.back()->d;
}

static BsPlan &get_bsplan(Ctx &c, int lmax, int mmax, int spin, int ncol)
{
  for (auto &p : c.bsplans)
    if (p->lmax == lmax && p->mmax == mmax && p->spin == spin && p->ncol == ncol) return *p;
  need_build("BF16 piece table");
  auto bp = std::make_shared<BsPlan>();
  bp->lmax = lmax;
  bp->mmax = mmax;
  bp->spin = spin;
  bp->ncol = ncol;
  std::vector<int64_t> base(mmax + 1);
  int64_t t = 0;
  for (int m = 0; m <= mmax; ++m) {
    base[m] = t;
    t += (int64_t)3 * ncol * bs_pairs(lmax, std::max(m, spin));
  }
  bp->total = t;
  bp->d = upload(base);
  c.bsplans.push_back(std::move(bp));
  return *c.bsplans.back();
}

static int64_t *get_almplan(Ctx &c, const sharp_alm_info *a, int mmax)
{
  std::vector<char> key;
  auto put = [&](const void *p, size_t n) {
    const char *ch = static_cast<const char *>(p);
    key.insert(key.end(), ch, ch + n);
  };
  put(&a->nm, sizeof(int));
  put(a->mval, sizeof(int) * a->nm);
  put(a->mvstart, sizeof(ptrdiff_t) * a->nm);
  for (auto &p : c.alms)
    if (p->key == key) return p->mvstart_by_m;
  need_build("a_lm layout");
  std::vector<int64_t> mvs(mmax + 1);
  for (int mi = 0; mi < a->nm; ++mi) mvs[a->mval[mi]] = a->mvstart[mi];
  auto ap = std::make_shared<AlmPlan>();
  ap->key = std::move(key);
  ap->mvstart_by_m = upload(mvs);
  c.alms.push_back(std::move(ap));
  return c.alms.back()->mvstart_by_m;
}

// ---------------------------------------------------------------------------
// job normalisation (libsharp/sharp.c:840-867) and checks
// ---------------------------------------------------------------------------
struct Job {
  int type, spin, nmaps, nalm, ntrans, flags, lmax, mmax, deriv;
  const sharp_geom_info *g;
  const sharp_alm_info *a;
  const double *lfac = nullptr;   // optional f_l applied on the a_lm side
};

static Job make_job(int type, int spin, const sharp_geom_info *g,
                    const sharp_alm_info *a, int ntrans, int flags)
{
  if (type < 0 || type > 4) SHP_FAIL("bad type");
  if (ntrans <= 0 || ntrans > 100) SHP_FAIL("bad number of simultaneous transforms");
  Job j;
  j.deriv = (type == SHARP_ALM2MAP_DERIV1);
  if (j.deriv) spin = 1;
  if (type == SHARP_MAP2ALM) flags |= SHARP_USE_WEIGHTS;
  if (type == SHARP_Yt) type = SHARP_MAP2ALM;
  if (type == SHARP_WY) { type = SHARP_ALM2MAP; flags |= SHARP_USE_WEIGHTS; }
  if (spin < 0 || spin > a->lmax) SHP_FAIL("bad spin");
  j.type = (type == SHARP_MAP2ALM) ? SHARP_MAP2ALM : SHARP_ALM2MAP;
  j.spin = spin;
  j.nmaps = (j.deriv || spin > 0) ? 2 : 1;
  j.nalm = j.deriv ? 1 : ((spin > 0) ? 2 : 1);
  j.ntrans = ntrans;
  j.flags = flags & ~SHARP_NVMAX;
  j.g = g;
  j.a = a;
  j.lmax = a->lmax;
  // every m from 0 to nm-1 exactly once (libsharp/sharp.c:238-251)
  std::vector<char> seen(a->nm, 0);
  for (int i = 0; i < a->nm; ++i) {
    const int mc = a->mval[i];
    if (mc < 0 || mc >= a->nm) SHP_FAIL("not all m values are present");
    if (seen[mc]) SHP_FAIL("duplicate m value");
    seen[mc] = 1;
  }
  j.mmax = a->nm - 1;
  if ((flags & SHARP_HIP_MIXED_FP32) && (flags & SHARP_HIP_MIXED_BF16X3))
    SHP_FAIL("SHARP_HIP_MIXED_FP32 and SHARP_HIP_MIXED_BF16X3 exclude each other");
  if (j.lmax < spin) SHP_FAIL("incorrect l_max: must be >= spin");
  if (j.lmax < j.mmax) SHP_FAIL("incorrect l_max: must be >= m_max");
  if (j.flags & SHARP_NO_FFT) {
    for (int i = 0; i < g->npairs; ++i) {
      if (g->pair[i].r1.nph != j.mmax + 1) SHP_FAIL("bad ring size");
      if (g->pair[i].r2.nph > 0 && g->pair[i].r2.nph != j.mmax + 1)
        SHP_FAIL("bad ring size");
    }
  }
  return j;
}

static unsigned long long op_count(const Job &j, const GeomPlan &gp,
                                   const GeomPlan::Mlim &ml)
{
  unsigned long long triples = 0;
  for (int k = 0; k < gp.nslots; ++k) {
    const long long M = std::min(ml.h[k], j.mmax);
    if (M < 0) continue;
    triples += (unsigned long long)((M + 1) * (long long)(j.lmax + 1) - M * (M + 1) / 2);
  }
  const unsigned long long per =
      j.deriv ? 12 + 8ull * j.ntrans
              : (j.spin == 0 ? 4 + 4ull * j.ntrans : 12 + 16ull * j.ntrans);
  return triples * per;
}

// runs fn() between two events on st when profiling is on
template <typename F>
static void timed(Ctx &c, hipStream_t st, const char *name, double flops,
                  double bytes, F &&fn)
{
  if (!c.prof.on) {
    fn();
    return;
  }
  Profiler::Pending p{c.prof.stage(name), c.prof.event(), c.prof.event(), flops, bytes};
  HIP_CHECK(hipEventRecord(p.a, st));
  fn();
  HIP_CHECK(hipEventRecord(p.b, st));
  c.prof.pending.push_back(p);
}

// ---------------------------------------------------------------------------
// device pipeline
// ---------------------------------------------------------------------------
// Everything one job needs on the device: its plans, the launch arguments of
// every stage and the algorithmic work of each stage (for the profiler).
struct JobState {
  Job j;
  // the ring FFT stage's extra streams, or nullptr: one stream
  const FftFork *fk = nullptr;
  GeomPlan *gp = nullptr;
  GeomPlan::Mlim *ml = nullptr;
  AlmArgs aa{};
  RingArgs ra{};
  LegArgs la{};
  FftClasses fc{};
  int ncomp_alm = 0, ncomp_map = 0;
  double triples = 0., c0 = 0., c1 = 0.;
  double ring_bytes_fft = 0., ring_bytes_all = 0., ring_bytes_dft = 0., alm_bytes = 0.;
  double ring_bytes_long = 0.;
  // scratch of the long-ring FFTs: lslots slots of 2 * 2^lg_max values
  double2 *lscratch = nullptr;
  int lslots = 0;
  // the cached plans the arguments above point into, held for as long as
  // the state lives (a batch outlives sharp_hip_release)
  std::vector<std::shared_ptr<void>> keep;
};

// the cache entry for which is(entry) holds
template <typename T, typename Pred>
static std::shared_ptr<void> owner(const std::vector<std::shared_ptr<T>> &v, Pred is)
{
  for (auto &e : v)
    if (is(*e)) return e;
  SHP_FAIL("internal error: plan not in the cache");
}

static size_t phase_elems(const Job &j, const GeomPlan &gp)
{
  const int nmb = (j.mmax + MB) / MB;
  return (size_t)j.ntrans * j.nmaps * nmb * gp.nslots * 2 * MB;
}

// kAlmSlackRows rows of slack: the mixed synthesis kernels read up to 15 rows
// past lmax (masked) without bounds checks
static size_t almtmp_elems(const Job &j)
{
  return ((size_t)tri_ofs(j.mmax + 1, j.lmax) + kAlmSlackRows) * j.ntrans * j.nalm;
}

// the mixed forms: FP32 matrix cores, or BF16 matrix cores on three-piece
// splits of the FP32 operands
static bool is_mixed(int flags)
{
  return (flags & (SHARP_HIP_MIXED_FP32 | SHARP_HIP_MIXED_BF16X3)) != 0;
}
static int mixed_contr(int flags)
{
  return (flags & SHARP_HIP_MIXED_BF16X3) ? kContrBf16x3 : kContrF32;
}

// norm_l * f_l on the device for a job with per-l factors, copied on `st`
// through pinned memory (reused once the previous copy out of it has ended)
static double *upload_lfac(Ctx &c, const Job &j, hipStream_t st)
{
  std::vector<double> n = norm_values(j.lmax, j.spin, j.deriv);
  const size_t bytes = n.size() * sizeof(double);
  if (c.lfac_ev) HIP_CHECK(hipEventSynchronize(c.lfac_ev));
  else HIP_CHECK(hipEventCreateWithFlags(&c.lfac_ev, hipEventDisableTiming));
  if (bytes > c.lfac_cap) {
    if (c.lfac_host) HIP_CHECK(hipHostFree(c.lfac_host));
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c.lfac_host), bytes));
    c.lfac_cap = bytes;
  }
  for (size_t l = 0; l < n.size(); ++l) c.lfac_host[l] = n[l] * j.lfac[l];
  double *d = static_cast<double *>(c.bufs(st).lfac.get(bytes, st, true));
  HIP_CHECK(hipMemcpyAsync(d, c.lfac_host, bytes, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipEventRecord(c.lfac_ev, st));
  return d;
}

static JobState setup_job(Ctx &c, const Job &j, void *const *dalm, void *const *dmap,
                          int64_t alm_base, int64_t map_base, double2 *phase,
                          double2 *almtmp, unsigned long long *opcnt, hipStream_t st)
{
  JobState s;
  s.j = j;
  GeomPlan &gp = get_geom(c, j.g);
  GeomPlan::Mlim &ml = get_mlim(gp, j.lmax, j.spin);
  YlmPlan &yp = get_ylm(c, j.lmax, j.mmax, j.spin);
  double *d_norm = get_norm(c, j.lmax, j.spin, j.deriv);
  int64_t *d_mvs = get_almplan(c, j.a, j.mmax);
  s.gp = &gp;
  s.ml = &ml;
  s.keep = {owner(c.geoms, [&](const GeomPlan &x) { return &x == &gp; }),
            owner(c.ylms, [&](const YlmPlan &x) { return &x == &yp; }),
            owner(c.norms, [&](const NormPlan &x) { return x.d == d_norm; }),
            owner(c.alms, [&](const AlmPlan &x) { return x.mvstart_by_m == d_mvs; })};
  const unsigned long long ops = op_count(j, gp, ml);
  if (opcnt) *opcnt = ops;

  const int ncomp_alm = j.ntrans * j.nalm, ncomp_map = j.ntrans * j.nmaps;
  s.ncomp_alm = ncomp_alm;
  s.ncomp_map = ncomp_map;
  const int nmb = (j.mmax + MB) / MB;

  AlmArgs &aa = s.aa;
  aa.lmax = j.lmax;
  aa.mmax = j.mmax;
  aa.ncomp = ncomp_alm;
  aa.flags = j.flags;
  aa.packed = j.a->flags & SHARP_PACKED;
  aa.stride = j.a->stride;
  aa.alm_base = alm_base;
  aa.mvstart_by_m = d_mvs;
  aa.norm_l = j.lfac ? upload_lfac(c, j, st) : d_norm;
  aa.almtmp = almtmp;
  // the mixed syntheses read FP32 a_lm
  aa.tmp_f32 = (is_mixed(j.flags) && j.type == SHARP_ALM2MAP) ? 1 : 0;
  for (int i = 0; i < ncomp_alm; ++i) aa.alms.p[i] = dalm[i];
  aa.bsplit = nullptr;
  aa.bs_base = nullptr;
  aa.bs_ncol = 0;
  aa.bs_deriv = j.deriv ? 1 : 0;
  aa.spin = j.spin;
  aa.part = nullptr;
  aa.gmax = ml.gmax;
  aa.ngroups = (int)ml.hgmax.size();
  aa.c0 = 0;
  aa.npc = ncomp_alm;
  aa.m0 = 0;
  aa.m1 = j.mmax + 1;
  aa.part_row0 = 0;
  aa.part_rows = tri_ofs(j.mmax + 1, j.lmax);

  RingArgs &ra = s.ra;
  ra.nslots = gp.nslots;
  ra.nmb = nmb;
  ra.mmax = j.mmax;
  ra.ncomp = ncomp_map;
  ra.flags = j.flags;
  ra.map_base = map_base;
  ra.ring = gp.ring;
  ra.phase = phase;
  ra.nphmax = gp.nphmax;
  ra.dp = (j.flags & SHARP_DP) ? 1 : 0;
  ra.max_lds = c.max_lds;
  for (int i = 0; i < ncomp_map; ++i) ra.maps.p[i] = dmap[i];

  LegArgs &la = s.la;
  la.lmax = j.lmax;
  la.mmax = j.mmax;
  la.nslots = gp.nslots;
  la.nmb = nmb;
  la.spin = j.spin;
  la.deriv1 = j.deriv;
  la.ncomp_alm = ncomp_alm;
  la.ncomp_map = ncomp_map;
  la.cth = gp.cth;
  la.sth = gp.sth;
  la.mlim = ml.d;
  la.ispair = gp.ispair;
  la.tab = yp.tab;
  if (is_mixed(j.flags)) {
    // the mixed kernels run the recurrence with a unit lambda_{l-2} factor;
    // every kernel reading this table computes lambda_l / s_l, and the gather
    // and scatter apply s_l
    ensure_unit_table(yp);
    la.tab = yp.tab1;
    la.unit_tab = 1;
    aa.lsc = yp.lsc;
  }
  la.mfac = yp.mfac;
  la.prefac = yp.prefac;
  la.fscale = yp.fscale;
  la.alm = almtmp;
  la.almout = almtmp;
  la.phase = phase;
  la.ngroups = (int)ml.hgmax.size();
  la.gmax = ml.gmax;
  la.m0 = 0;
  la.m1 = j.mmax + 1;
  la.part_row0 = 0;
  la.part_rows = tri_ofs(j.mmax + 1, j.lmax);
  la.bsplit = nullptr;
  la.bs_base = nullptr;
  la.bs_ncol = 0;

  // triples (m, ring pair, l) with m <= mlim of the pair, and the flop model
  // constants of libsharp/sharp_core_inc2.c:171-173,440-442,583-585
  s.c0 = j.spin == 0 ? 4. : 12.;
  s.c1 = j.deriv ? 8. : (j.spin == 0 ? 4. : 16.);
  s.triples = (double)ops / (s.c0 + s.c1 * j.ntrans);

  int64_t nring_fft = 0;
  s.fc.btab = gp.btab;
  for (int cl = 0; cl < kFftKinds * kNumFftClasses; ++cl) {
    s.fc.plans[cl] = gp.fft[cl];
    s.fc.hplans[cl] = gp.hfft[cl].data();
    s.fc.count[cl] = gp.nfft[cl];
    nring_fft += gp.nfft[cl];
  }
  const double esz = (j.flags & SHARP_DP) ? 8. : 4.;
  // compulsory traffic of the ring stage: the phases of every ring and its pixels
  s.ring_bytes_fft =
      (double)ncomp_map * ((double)nring_fft * (j.mmax + 1) * 16. + (double)gp.npix_fft * esz);
  s.ring_bytes_all =
      (double)ncomp_map * ((double)gp.nall * (j.mmax + 1) * 16. + (double)gp.npix_all * esz);
  s.ring_bytes_long =
      (double)ncomp_map * ((double)gp.nlong * (j.mmax + 1) * 16. + (double)gp.npix_long * esz);
  s.ring_bytes_dft = std::max(0., s.ring_bytes_all - s.ring_bytes_fft - s.ring_bytes_long);
  s.alm_bytes = (double)tri_ofs(j.mmax + 1, j.lmax) * ncomp_alm * (16. + 2. * esz);
  return s;
}

// SHARP_HIP_RING_DFT=1 routes every ring through the direct-summation kernels
// (a cross-check of the FFT path; slow for long rings)
static bool force_ring_dft()
{
  static const bool f = [] {
    const char *e = std::getenv("SHARP_HIP_RING_DFT");
    return e && e[0] == '1';
  }();
  return f;
}

// whether the rings go through the direct sums only
static bool rings_direct(const Job &j) { return (j.flags & SHARP_NO_FFT) || force_ring_dft(); }

// workgroups (and scratch slots) of the long-ring FFTs of a job
static int long_slots(const GeomPlan &gp, const Job &j, int cap = 2048)
{
  if (!gp.nlong || rings_direct(j)) return 0;
  return std::min(gp.nlong * j.ntrans * j.nmaps, cap);
}
static size_t long_scratch_bytes(const GeomPlan &gp, int slots)
{
  return (size_t)slots * 2 * ((size_t)1 << gp.lg_max) * sizeof(double2);
}

static LongRingArgs long_args(const JobState &s)
{
  const GeomPlan &gp = *s.gp;
  return LongRingArgs{gp.longp, gp.nlong, gp.ltab, s.lscratch, 2 * ((int64_t)1 << gp.lg_max)};
}

static void stage_rings_to_phase(Ctx &c, JobState &s, hipStream_t st)
{
  const GeomPlan &gp = *s.gp;
  if (rings_direct(s.j)) {
    timed(c, st, "ring_dft_anal", 0., s.ring_bytes_all,
          [&] { launch_ring2phase_dft(s.ra, gp.all_list, gp.nall, st); });
  } else {
    timed(c, st, "ring_fft_anal", 0., s.ring_bytes_fft,
          [&] { launch_ring_fft(false, s.ra, s.fc, st, s.fk); });
    if (gp.nlong)
      timed(c, st, "ring_long_anal", 0., s.ring_bytes_long,
            [&] { launch_long_rings(false, s.ra, long_args(s), s.lslots, st); });
    if (gp.ndft)
      timed(c, st, "ring_dft_anal", 0., s.ring_bytes_dft,
            [&] { launch_ring2phase_dft(s.ra, gp.dft_list, gp.ndft, st); });
  }
}

static void stage_phase_to_rings(Ctx &c, JobState &s, hipStream_t st)
{
  const GeomPlan &gp = *s.gp;
  if (rings_direct(s.j)) {
    timed(c, st, "ring_dft_synth", 0., s.ring_bytes_all,
          [&] { launch_phase2ring_dft(s.ra, gp.all_list, gp.nall, st); });
  } else {
    timed(c, st, "ring_fft_synth", 0., s.ring_bytes_fft,
          [&] { launch_ring_fft(true, s.ra, s.fc, st, s.fk); });
    if (gp.nlong)
      timed(c, st, "ring_long_synth", 0., s.ring_bytes_long,
            [&] { launch_long_rings(true, s.ra, long_args(s), s.lslots, st); });
    if (gp.ndft)
      timed(c, st, "ring_dft_synth", 0., s.ring_bytes_dft,
            [&] { launch_phase2ring_dft(s.ra, gp.dft_list, gp.ndft, st); });
  }
}

static void stage_gather(Ctx &c, JobState &s, hipStream_t st)
{
  timed(c, st, "alm_gather", 0., s.alm_bytes, [&] { launch_alm_gather(s.aa, st); });
}

// FP64 Legendre stage of one job
static void stage_legendre_fp64(Ctx &c, JobState &s, hipStream_t st)
{
  const Job &j = s.j;
  const int gmax = (j.spin == 0) ? 4 : 2;
  const char *syn_name = j.deriv ? "legendre_synth_deriv1"
                                 : (j.spin == 0 ? "legendre_synth_spin0" : "legendre_synth_spin");
  const char *ana_name = j.spin == 0 ? "legendre_anal_spin0" : "legendre_anal_spin";
  int t = 0;
  while (t < j.ntrans) {
    int g = gmax;
    while (g > j.ntrans - t) g >>= 1;
    s.la.t0 = t;
    if (j.type == SHARP_ALM2MAP)
      timed(c, st, syn_name, s.triples * (s.c0 + s.c1 * g), 0., [&] { launch_synth(s.la, g, st); });
    else
      timed(c, st, ana_name, s.triples * (s.c0 + s.c1 * g), 0., [&] { launch_anal(s.la, g, st); });
    t += g;
  }
}


static const char *mixed_name(const Job &j)
{
  const bool bf = (j.flags & SHARP_HIP_MIXED_BF16X3) != 0;
  if (j.type == SHARP_ALM2MAP) {
    if (j.deriv) return bf ? "bf16x3_synth_deriv1" : "mixed_synth_deriv1";
    if (j.spin == 0) return bf ? "bf16x3_synth_spin0" : "mixed_synth_spin0";
    return bf ? "bf16x3_synth_spin" : "mixed_synth_spin";
  }
  if (j.spin == 0) return bf ? "bf16x3_anal_spin0" : "mixed_anal_spin0";
  return bf ? "bf16x3_anal_spin" : "mixed_anal_spin";
}

// SHARP_HIP_FP64_SYNTHESIS=valu selects the FP64 synthesis kernels that
// accumulate on the vector ALUs (legendre.hip, up to 4 spin-0 / 2 spin-s
// transforms per launch) instead of the FP64 matrix-core kernels
static bool fp64_synth_valu()
{
  const char *e = std::getenv("SHARP_HIP_FP64_SYNTHESIS");
  return e && std::strcmp(e, "valu") == 0;
}

// SHARP_HIP_FP64_ANALYSIS=reduce selects the FP64 analysis kernels that sum
// over ring pairs with cross-lane reductions (legendre.hip) instead of the
// FP64 matrix-core kernels; both are exact to FP64 rounding
static bool fp64_anal_reduce()
{
  const char *e = std::getenv("SHARP_HIP_FP64_ANALYSIS");
  return e && std::strcmp(e, "reduce") == 0;
}

// transforms per analysis launch of the matrix-core kernels
static int anal_tmax(const Job &j)
{
  return is_mixed(j.flags) ? ((j.spin == 0) ? kMixMaxT0 : kMixMaxTS)
                           : ((j.spin == 0) ? kF64MaxT0 : kF64MaxTS);
}

// The partial sums of one analysis launch are capped at SHARP_HIP_PART_CAP_MB
// (default 32 GiB; 16 GiB at HEALPix Nside 2048, lmax 4096, 8 spin-2
// transforms): larger analyses run in chunks of orders m, each chunk's
// Legendre launch followed by its scatter.
static size_t part_cap()
{
  static const size_t cap = [] {
    const char *e = std::getenv("SHARP_HIP_PART_CAP_MB");
    const long v = e ? std::atol(e) : 0;
    return v > 0 ? (size_t)v << 20 : (size_t)32 << 30;
  }();
  return cap;
}

// bytes of the partial sums of one canonical row over ng ring groups
static size_t part_row_bytes(const Job &j, int ng)
{
  const size_t esz = is_mixed(j.flags) ? sizeof(float) : sizeof(double);
  const int npc = std::min(anal_tmax(j), j.ntrans) * j.nalm;
  return (size_t)ng * 2 * npc * esz;
}

// boundaries m_0 = 0 < m_1 < ... = mmax+1 (multiples of 8 but the last) of
// the analysis chunks whose partial sums fit the cap (at least 8 orders m
// per chunk)
static std::vector<int> part_chunks(const Job &j, int ng)
{
  const size_t per_row = part_row_bytes(j, ng), cap = part_cap();
  std::vector<int> b{0};
  int m = 0;
  while (m <= j.mmax) {
    int e = m;
    size_t rows = 0;
    while (e <= j.mmax) {
      const int e2 = std::min(e + 8, j.mmax + 1);
      const size_t add = (size_t)(tri_ofs(e2, j.lmax) - tri_ofs(e, j.lmax));
      if (e > m && (rows + add) * per_row > cap) break;
      rows += add;
      e = e2;
    }
    b.push_back(e);
    m = e;
  }
  return b;
}

// bytes of the per-group partial sums of the largest analysis launch (ng
// groups of ring pairs), 0 when the job needs none
static size_t part_bytes(const Job &j, int ng)
{
  if (j.type == SHARP_ALM2MAP || (!is_mixed(j.flags) && fp64_anal_reduce())) return 0;
  const std::vector<int> b = part_chunks(j, ng);
  size_t most = 0;
  for (size_t i = 0; i + 1 < b.size(); ++i)
    most = std::max(most, (size_t)(tri_ofs(b[i + 1], j.lmax) - tri_ofs(b[i], j.lmax)));
  return most * part_row_bytes(j, ng);
}

// Analysis Legendre stage and a_lm scatter of one job, in transform groups:
// the matrix-core kernels write per-group partial sums into `part` (of
// part_bytes(j, groups) bytes; FP32 for the mixed form, FP64 for the FP64
// form) and the scatter adds them.
static void stage_analysis(Ctx &c, JobState &s, void *part, hipStream_t st)
{
  const Job &j = s.j;
  const bool mixed = is_mixed(j.flags);
  if (!mixed && fp64_anal_reduce()) {
    stage_legendre_fp64(c, s, st);
    timed(c, st, "alm_scatter", 0., s.alm_bytes, [&] { launch_alm_scatter(s.aa, st); });
    return;
  }
  const int tmax = anal_tmax(j);
  const char *name = mixed ? mixed_name(j)
                           : (j.spin == 0 ? "legendre_anal_spin0" : "legendre_anal_spin");
  LegArgs &la = s.la;
  const std::vector<int> mb = part_chunks(j, la.ngroups);
  const double rows = (double)tri_ofs(j.mmax + 1, j.lmax);
  for (int t = 0; t < j.ntrans; t += tmax) {
    la.t0 = t;
    la.ntl = std::min(tmax, j.ntrans - t);
    const int npc = la.ntl * j.nalm;
    la.part = part;
    la.part_ld = 2 * npc;
    la.part_col0 = 0;
    s.aa.part = la.part;
    s.aa.part_f64 = mixed ? 0 : 1;
    s.aa.part_ld = la.part_ld;
    s.aa.part_c0 = t * j.nalm;
    s.aa.c0 = t * j.nalm;
    s.aa.npc = npc;
    for (size_t k = 0; k + 1 < mb.size(); ++k) {
      // the chunk's orders m, its rows at the start of the partial buffer
      la.m0 = s.aa.m0 = mb[k];
      la.m1 = s.aa.m1 = mb[k + 1];
      la.part_row0 = s.aa.part_row0 = tri_ofs(mb[k], j.lmax);
      la.part_rows = s.aa.part_rows = tri_ofs(mb[k + 1], j.lmax) - la.part_row0;
      const double share = (double)la.part_rows / rows;
      timed(c, st, name, share * s.triples * (s.c0 + s.c1 * la.ntl), 0.,
            [&] { launch_manal(la, mixed ? mixed_contr(j.flags) : kContrF64, st); });
      timed(c, st, "alm_scatter", 0., share * s.alm_bytes * npc / s.ncomp_alm,
            [&] { launch_alm_scatter(s.aa, st); });
    }
  }
  // back to the whole range for any later use of the arguments
  la.m0 = s.aa.m0 = 0;
  la.m1 = s.aa.m1 = j.mmax + 1;
  la.part_row0 = s.aa.part_row0 = 0;
  la.part_rows = s.aa.part_rows = tri_ofs(j.mmax + 1, j.lmax);
}

// Work memory of one call supplied by the caller of run_device instead of
// the per-stream buffers (sharp_hip_execute_device_noalloc).
struct WorkPtrs {
  double2 *phase = nullptr, *almtmp = nullptr;
  double2 *lscratch = nullptr;   // long-ring FFT slots (lslots of them)
  int lslots = 0;
  void *part = nullptr;          // part_bytes
  unsigned *bsplit = nullptr;    // bs_words: the BF16 pieces of the a_lm
  int64_t *bs_base = nullptr;    // [mmax+1] word offsets (filled on the device)
};

// BF16 spin-s synthesis: its a_lm are split into BF16 pieces by the gather
static bool uses_bsplit(const Job &j)
{
  return is_mixed(j.flags) && mixed_contr(j.flags) == kContrBf16x3 && j.spin > 0 &&
         j.type == SHARP_ALM2MAP;
}
static int bs_ncol(const Job &j) { return 4 * j.ntrans + 1; }
// 32-bit words of the BF16 pieces (the layout of bs_pairs)
static int64_t bs_words(const Job &j)
{
  int64_t t = 0;
  for (int m = 0; m <= j.mmax; ++m)
    t += (int64_t)3 * bs_ncol(j) * bs_pairs(j.lmax, std::max(m, j.spin));
  return t;
}

// The extra streams of the ring FFT stage on launch stream st, created with
// st's device current on first use; nullptr (one stream) while st is being
// captured into a graph.
static const FftFork *ring_fork(Ctx::StreamBufs &b, hipStream_t st)
{
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_CHECK(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return nullptr;
  if (!b.has_fork) {
    for (int i = 0; i < 2; ++i) {
      HIP_CHECK(hipStreamCreateWithFlags(&b.fork.aux[i], hipStreamNonBlocking));
      HIP_CHECK(hipEventCreateWithFlags(&b.fork.join[i], hipEventDisableTiming));
    }
    HIP_CHECK(hipEventCreateWithFlags(&b.fork.fork, hipEventDisableTiming));
    b.has_fork = true;
  }
  return &b.fork;
}

static void run_device(Ctx &c, const Job &j, void *const *dalm, void *const *dmap,
                       int64_t alm_base, int64_t map_base, hipStream_t st,
                       unsigned long long *opcnt, const WorkPtrs *w = nullptr)
{
  GeomPlan &gp = get_geom(c, j.g);
  Ctx::StreamBufs *sb = w ? nullptr : &c.bufs(st);
  double2 *phase = w ? w->phase
                     : static_cast<double2 *>(
                           sb->phase.get(phase_elems(j, gp) * sizeof(double2), st, true));
  double2 *almtmp = w ? w->almtmp
                      : static_cast<double2 *>(
                            sb->almtmp.get(almtmp_elems(j) * sizeof(double2), st, true));
  JobState s = setup_job(c, j, dalm, dmap, alm_base, map_base, phase, almtmp, opcnt, st);
  if (!w) s.fk = ring_fork(*sb, st);
  if (w) {
    s.lslots = std::min(w->lslots, long_slots(gp, j));
    s.lscratch = w->lscratch;
  } else {
    s.lslots = long_slots(gp, j);
    if (s.lslots)
      s.lscratch = static_cast<double2 *>(
          sb->lscratch.get(long_scratch_bytes(gp, s.lslots), st, true));
  }

  const bool mixed = is_mixed(j.flags);
  if (j.type != SHARP_ALM2MAP) {
    stage_rings_to_phase(c, s, st);
    void *part = w ? w->part : sb->part.get(part_bytes(j, (int)s.ml->hgmax.size()), st, true);
    stage_analysis(c, s, part, st);
    return;
  }
  if (uses_bsplit(j)) {
    const int ncol = bs_ncol(j);
    unsigned *bs;
    const int64_t *base;
    if (w) {
      bs = w->bsplit;
      launch_bs_base(w->bs_base, j.lmax, j.mmax, j.spin, ncol, st);
      base = w->bs_base;
    } else {
      BsPlan &bp = get_bsplan(c, j.lmax, j.mmax, j.spin, ncol);
      bs = static_cast<unsigned *>(sb->bsplit.get((size_t)bp.total * sizeof(unsigned), st, true));
      base = bp.d;
    }
    s.aa.bsplit = bs;
    s.aa.bs_base = base;
    s.aa.bs_ncol = ncol;
    s.la.bsplit = bs;
    s.la.bs_base = base;
    s.la.bs_ncol = ncol;
  }
  if (s.aa.bsplit)
    timed(c, st, "alm_gather", 0., s.alm_bytes, [&] { launch_alm_gather_split(s.aa, st); });
  else
    stage_gather(c, s, st);
  if (!mixed && !fp64_synth_valu()) {
    // FP64 matrix-core syntheses: up to 16 (spin 0) or 8 (spin s) transforms
    // per launch, the recurrence once for all of them
    const int tmax = (j.spin == 0) ? kF64SynT0 : kF64SynTS;
    const char *name = j.deriv ? "legendre_synth_deriv1"
                               : (j.spin == 0 ? "legendre_synth_spin0" : "legendre_synth_spin");
    LegArgs &la = s.la;
    for (int t = 0; t < j.ntrans; t += tmax) {
      la.t0 = t;
      la.ntl = std::min(tmax, j.ntrans - t);
      timed(c, st, name, s.triples * (s.c0 + s.c1 * la.ntl), 0.,
            [&] { launch_synth64(la, st); });
    }
  } else if (!mixed) {
    stage_legendre_fp64(c, s, st);
  } else {
    // mixed precision: transforms in groups of up to 16 (spin 0) or 8 (spin
    // s) per launch, i.e. up to 32 real MFMA columns
    const int tmax = (j.spin == 0) ? kMixMaxT0 : kMixMaxTS;
    LegArgs &la = s.la;
    for (int t = 0; t < j.ntrans; t += tmax) {
      la.t0 = t;
      la.ntl = std::min(tmax, j.ntrans - t);
      timed(c, st, mixed_name(j), s.triples * (s.c0 + s.c1 * la.ntl), 0.,
            [&] { launch_msynth(la, mixed_contr(j.flags), st); });
    }
  }
  stage_phase_to_rings(c, s, st);
}

// extent [lo, hi] (in real units) touched by an a_lm array
static void alm_extent(const sharp_alm_info *a, int64_t &lo, int64_t &hi)
{
  lo = INT64_MAX;
  hi = INT64_MIN;
  const bool packed = a->flags & SHARP_PACKED;
  for (int mi = 0; mi < a->nm; ++mi) {
    const int m = a->mval[mi];
    const bool realonly = packed && m == 0;
    int64_t ofs = a->mvstart[mi], stride = a->stride;
    if (!packed) ofs *= 2;
    if (!realonly) stride *= 2;
    const int64_t e0 = ofs + (int64_t)m * stride, e1 = ofs + (int64_t)a->lmax * stride;
    const int64_t add = realonly ? 0 : 1;
    lo = std::min(lo, std::min(e0, e1));
    hi = std::max(hi, std::max(e0, e1) + add);
  }
}

// extent [lo, hi] (in pixel units) touched by a map
static void map_extent(const sharp_geom_info *g, int64_t &lo, int64_t &hi)
{
  lo = INT64_MAX;
  hi = INT64_MIN;
  for (int i = 0; i < g->npairs; ++i)
    for (const sharp_ringinfo *r : {&g->pair[i].r1, &g->pair[i].r2}) {
      if (r->nph <= 0) continue;
      const int64_t e0 = r->ofs, e1 = r->ofs + (int64_t)(r->nph - 1) * r->stride;
      lo = std::min(lo, std::min(e0, e1));
      hi = std::max(hi, std::max(e0, e1));
    }
}

// ---------------------------------------------------------------------------
// batches of transforms of uneven size: one persistent launch per Legendre
// kernel kind over the tiles of every job, recorded into a HIP graph
// ---------------------------------------------------------------------------
struct BatchWork {
  int kind = 0, ncb = 1, grid = 0, nitems = 0;
  int flat_grid = 0;    // workgroups of the flat synthesis launch
  int tg = 0;           // transforms of every item (FP64 syntheses), else 0
  BatchItem *items = nullptr;
  int *xofs = nullptr;
  int *ctr = nullptr;   // 8 item counters, one per XCD part of the list
  double flops = 0.;
  const char *name = "";
};

// The tile shape of a batch, chosen by batch_tune.
struct BatchTiles {
  int tgroup = 0;      // transforms per tile: the form's maximum >> tgroup
  int wg_per_cu = 2;   // persistent workgroups per CU
  int flat = 0;        // syntheses: 1 = one flat launch (a wavefront per tile)
                       // instead of persistent workgroups
  int shape = 0;       // wavefronts per workgroup: synthesis m-blocks of 8 (0) or 4
                       // (1) orders m, one per wavefront, of 64 ring pairs;
                       // analysis ring block of 4 (0) or 2 (1) wavefronts of 64
                       // ring pairs
};

struct BatchPlan {
  int device = 0;
  bool synth = true;
  int contr = kContrF64;   // Legendre contraction: FP64, FP32 or BF16-split matrix cores
  std::vector<Job> jobs;
  std::vector<std::vector<void *>> almp, mapp;
  std::vector<DevBuf> phase, almtmp, part, bsplit, gmax2;
  DevBuf lscratch;   // long-ring FFT slots, shared by the jobs (run in turn)
  std::vector<JobState> st;
  // per-job launch arguments for each analysis ring block (4, 2 wavefronts):
  // their ring groups and group maxima of mlim differ
  LegArgs *djobs[2] = {nullptr, nullptr};
  std::vector<int> ngroups[2];
  std::vector<int *> dgmax[2];
  std::vector<BatchWork> work;
  BatchTiles tiles;
  bool use_graph = true;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void free_work()
  {
    for (auto &w : work) { (void)hipFree(w.items); (void)hipFree(w.xofs); (void)hipFree(w.ctr); }
    work.clear();
  }
  ~BatchPlan()
  {
    free_work();
    for (auto *d : djobs) (void)hipFree(d);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    for (auto *v : {&phase, &almtmp, &part, &bsplit, &gmax2})
      for (auto &b : *v) b.release();
    lscratch.release();
  }
};

static int batch_kind(const Job &j)
{
  if (j.type == SHARP_ALM2MAP) return j.deriv ? kBatchSynthD : (j.spin == 0 ? kBatchSynth0 : kBatchSynthS);
  return j.spin == 0 ? kBatchAnal0 : kBatchAnalS;
}

static const char *batch_name(int kind)
{
  switch (kind) {
  case kBatchSynth0: return "batch_synth_spin0";
  case kBatchSynthS: return "batch_synth_spin";
  case kBatchSynthD: return "batch_synth_deriv1";
  case kBatchAnal0: return "batch_anal_spin0";
  default: return "batch_anal_spin";
  }
}

// largest transform group of a kernel kind in a form: the mixed forms take
// up to 32 real MFMA columns; the FP64 matrix-core analysis 16; the FP64
// VALU syntheses 4 (spin 0) or 2 (spin s) transforms per recurrence
static int batch_tmax(int kind, int contr)
{
  const bool spin0 = (kind == kBatchSynth0 || kind == kBatchAnal0);
  const bool synth = kind <= kBatchSynthD;
  if (contr != kContrF64) return spin0 ? kMixMaxT0 : kMixMaxTS;
  // the FP64 matrix-core syntheses take the single-job groups (two blocks of
  // 16 real columns); the vector-ALU ones (SHARP_HIP_FP64_SYNTHESIS=valu) 4 / 2
  if (synth) return fp64_synth_valu() ? (spin0 ? 4 : 2) : (spin0 ? kF64SynT0 : kF64SynTS);
  // the batch FP64 analysis kernels hold one block of 16 columns
  return spin0 ? kF64MaxT0 : 4;
}

// analysis wavefronts per workgroup of a ring-block choice
static int batch_waves(int shape) { return shape ? 2 : 4; }

template <typename G> static size_t nitems_kind(const std::vector<G> &groups)
{
  size_t n = 0;
  for (auto &g : groups) n += g.items.size();
  return n;
}

// Work lists for a tile shape: tiles grouped by (job, transform group,
// m-block or m) so that a group's tiles run on one XCD; groups are dealt to
// the 8 XCDs largest first onto the least loaded one.
static void build_work(BatchPlan &b, int ncu)
{
  b.free_work();
  const BatchTiles &tl = b.tiles;
  for (int kind = kBatchSynth0; kind <= kBatchAnalS; ++kind) {
    const bool spin0 = (kind == kBatchSynth0 || kind == kBatchAnal0);
    const bool synth = kind <= kBatchSynthD;
    const int tg = std::max(1, batch_tmax(kind, b.contr) >> tl.tgroup);
    const int wi = synth ? 0 : tl.shape;
    // tl: the transforms of the group's items when the kind's kernel is
    // specialised by it (FP64 syntheses), else 0
    struct Group { double cost; std::vector<BatchItem> items; int tl = 0; };
    std::vector<Group> groups;
    double flops = 0.;
    for (size_t ji = 0; ji < b.jobs.size(); ++ji) {
      const Job &j = b.jobs[ji];
      if (batch_kind(j) != kind) continue;
      const JobState &s = b.st[ji];
      flops += s.triples * (s.c0 + s.c1 * j.ntrans);
      const int nslots = s.gp->nslots;
      int t0 = 0;
      while (t0 < j.ntrans) {
        int ntl = std::min(tg, j.ntrans - t0);
        // the vector-ALU FP64 syntheses have kernels for power-of-two groups
        // only
        const bool valu64 = synth && b.contr == kContrF64 && fp64_synth_valu();
        if (valu64)
          while (ntl & (ntl - 1)) ntl &= ntl - 1;
        if (synth) {
          // the 8 orders m of a phase block (whose phases share cache lines)
          // in one group, tiles of `waves` orders m x 64 slots
          const int waves = 2 * batch_waves(tl.shape), per = MB / waves;
          const int nmb = (j.mmax + MB) / MB, nsb = (nslots + 63) / 64;
          for (int mb = 0; mb < nmb; ++mb) {
            Group g;
            g.tl = valu64 ? ntl : 0;
            g.cost = (double)(j.lmax - mb * MB + 1) * MB * nsb * (s.c0 + s.c1 * ntl);
            for (int sb = 0; sb < nsb; ++sb)
              for (int h = 0; h < per; ++h)
                g.items.push_back(BatchItem{(int)ji, t0, ntl, 0, mb * per + h, sb, 0, 0});
            groups.push_back(std::move(g));
          }
        } else {
          const int gs = 64 * batch_waves(wi);
          const int ng = (nslots + gs - 1) / gs;
          std::vector<int> gm(ng, -1);
          for (int k = 0; k < nslots; ++k) gm[k / gs] = std::max(gm[k / gs], s.ml->h[k]);
          for (int m = 0; m <= j.mmax; ++m) {
            Group g;
            for (int gi = 0; gi < ng; ++gi)
              if (gm[gi] >= m)
                g.items.push_back(BatchItem{(int)ji, t0, ntl, 2 * t0 * j.nalm, m, gi, 0, 0});
            if (g.items.empty()) continue;
            g.cost = (double)(j.lmax - m + 1) * g.items.size() * gs * (s.c0 + s.c1 * ntl);
            groups.push_back(std::move(g));
          }
        }
        t0 += ntl;
      }
    }
    if (groups.empty()) continue;
    std::stable_sort(groups.begin(), groups.end(),
                     [](const Group &x, const Group &y) { return x.cost > y.cost; });
    std::vector<int> tls;
    for (auto &g : groups)
      if (std::find(tls.begin(), tls.end(), g.tl) == tls.end()) tls.push_back(g.tl);
    for (int tlv : tls) {
      std::vector<BatchItem> lists[8];
      double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (auto &g : groups) {
        if (g.tl != tlv) continue;
        const int x = (int)(std::min_element(load, load + 8) - load);
        load[x] += g.cost;
        lists[x].insert(lists[x].end(), g.items.begin(), g.items.end());
      }
      std::vector<BatchItem> all;
      std::vector<int> xofs(9, 0);
      for (int x = 0; x < 8; ++x) {
        xofs[x] = (int)all.size();
        all.insert(all.end(), lists[x].begin(), lists[x].end());
      }
      xofs[8] = (int)all.size();
      int maxntl = 1;
      for (const BatchItem &it : all) maxntl = std::max(maxntl, it.ntl);
      BatchWork w;
      w.kind = kind;
      w.tg = tlv;
      // blocks of 16 real columns: the mixed kernels and the FP64 matrix-core
      // syntheses take one or two
      const bool blocks = b.contr != kContrF64 || (synth && tlv == 0);
      w.ncb = (blocks && maxntl * (spin0 ? 2 : 4) > 16) ? 2 : 1;
      w.nitems = (int)all.size();
      w.grid = std::max(8, ((tl.wg_per_cu * ncu + 7) / 8) * 8);
      if (synth) {
        // flat launch: a workgroup per item, 8 parts
        int most = 0;
        for (int x = 0; x < 8; ++x) most = std::max(most, xofs[x + 1] - xofs[x]);
        w.flat_grid = 8 * most;
      }
      w.items = upload(all);
      w.xofs = upload(xofs);
      w.ctr = upload(std::vector<int>(8, 0));
      w.flops = flops * (double)all.size() / std::max<size_t>(1, nitems_kind(groups));
      w.name = batch_name(kind);
      b.work.push_back(w);
    }
  }
}

// the tile shape's launch arguments in every job's scatter
static void batch_scatter_args(BatchPlan &b)
{
  const int wi = b.tiles.shape;
  for (size_t i = 0; i < b.st.size(); ++i) {
    JobState &s = b.st[i];
    s.aa.gmax = b.dgmax[wi][i];
    s.aa.ngroups = b.ngroups[wi][i];
  }
}

static BatchLaunch batch_launch(const BatchPlan &b, const BatchWork &w, int shape)
{
  BatchLaunch l;
  l.kind = w.kind;
  l.ncb = w.ncb;
  l.contr = b.contr;
  l.shape = shape;
  l.tg = w.tg;
  l.flat = b.synth && b.tiles.flat;
  l.grid = w.grid;
  l.flat_grid = w.flat_grid;
  l.items = w.items;
  l.xofs = w.xofs;
  l.ctr = w.ctr;
  return l;
}

static void batch_run(Ctx &c, BatchPlan &b, hipStream_t st)
{
  const size_t n = b.jobs.size();
  const int shape = batch_waves(b.tiles.shape) * (b.synth ? 2 : 1);
  const LegArgs *dj = b.djobs[b.synth ? 0 : b.tiles.shape];
  if (b.synth) {
    for (size_t i = 0; i < n; ++i) {
      JobState &s = b.st[i];
      if (s.aa.bsplit)
        timed(c, st, "alm_gather", 0., s.alm_bytes, [&] { launch_alm_gather_split(s.aa, st); });
      else
        stage_gather(c, s, st);
    }
    for (auto &w : b.work)
      timed(c, st, w.name, w.flops, 0., [&] { launch_batch(batch_launch(b, w, shape), dj, st); });
    for (size_t i = 0; i < n; ++i) stage_phase_to_rings(c, b.st[i], st);
  } else {
    for (size_t i = 0; i < n; ++i) stage_rings_to_phase(c, b.st[i], st);
    for (auto &w : b.work)
      timed(c, st, w.name, w.flops, 0., [&] { launch_batch(batch_launch(b, w, shape), dj, st); });
    batch_scatter_args(b);
    for (size_t i = 0; i < n; ++i) {
      JobState &s = b.st[i];
      timed(c, st, "alm_scatter", 0., s.alm_bytes, [&] { launch_alm_scatter(s.aa, st); });
    }
  }
}

// key of a batch for the tuning cache
static std::string batch_key(const BatchPlan &b)
{
  std::string k = b.synth ? "S" : "A";
  k += std::to_string(b.contr);
  char buf[96];
  for (const Job &j : b.jobs) {
    const GeomPlan &gp = *b.st[&j - &b.jobs[0]].gp;
    unsigned long h = 1469598103934665603ul;
    for (char ch : gp.key) h = (h ^ (unsigned char)ch) * 1099511628211ul;
    snprintf(buf, sizeof(buf), "|%lx:%d:%d:%d:%d:%d", h, j.lmax, j.mmax, j.spin, j.deriv, j.ntrans);
    k += buf;
  }
  return k;
}

static std::map<std::string, BatchTiles> g_tune;

static const char *tune_file()
{
  const char *f = std::getenv("SHARP_HIP_TUNE_CACHE");
  return (f && f[0]) ? f : nullptr;
}

// cache file lines: "<key> <tgroup> <wg_per_cu> <shape> <flat>"
static bool tune_lookup(const std::string &key, BatchTiles &v)
{
  auto it = g_tune.find(key);
  if (it != g_tune.end()) { v = it->second; return true; }
  const char *f = tune_file();
  if (!f) return false;
  FILE *fp = fopen(f, "r");
  if (!fp) return false;
  std::vector<char> line(1 << 16);
  bool found = false;
  while (fgets(line.data(), (int)line.size(), fp)) {
    char *kend = strchr(line.data(), ' ');
    if (!kend) continue;
    *kend = 0;
    BatchTiles t;
    if (sscanf(kend + 1, "%d %d %d %d", &t.tgroup, &t.wg_per_cu, &t.shape, &t.flat) != 4)
      continue;
    if (key == line.data() && t.tgroup >= 0 && t.tgroup <= 2 && t.wg_per_cu >= 1 &&
        t.wg_per_cu <= 8 && (t.shape == 0 || t.shape == 1) && (t.flat == 0 || t.flat == 1)) {
      v = t;
      found = true;
    }
  }
  fclose(fp);
  if (found) g_tune[key] = v;
  return found;
}

static void tune_store(const std::string &key, const BatchTiles &v)
{
  g_tune[key] = v;
  const char *f = tune_file();
  if (!f) return;
  if (FILE *fp = fopen(f, "a")) {
    fprintf(fp, "%s %d %d %d %d\n", key.c_str(), v.tgroup, v.wg_per_cu, v.shape, v.flat);
    fclose(fp);
  }
}

// Project autotuner: times every tile shape -- transform group (the form's
// maximum, half, quarter) x dispatch (persistent workgroups per CU 1, 2, 4,
// fewer where fewer stay resident; syntheses also one flat launch with a
// workgroup per tile) x workgroup shape (synthesis: m-blocks of 8 or 4
// orders m, a wavefront of 64 ring pairs each; analysis: ring blocks of 4 or
// 2 wavefronts of 64 ring pairs) -- on the batch itself (one untimed and one
// timed run each, HIP events on the plan's stream) and keeps the fastest;
// the choice is cached by the batch's shape (and in SHARP_HIP_TUNE_CACHE,
// when set).  Shapes that give the same work lists as an earlier one (a
// transform group larger than every job's count) are skipped.
static void batch_tune(Ctx &c, BatchPlan &b, int ncu)
{
  BatchTiles best;
  const std::string key = batch_key(b);
  const char *e = std::getenv("SHARP_HIP_TUNE");
  const bool off = e && e[0] == '0';
  if (!tune_lookup(key, best) && !off) {
    float tbest = 1e30f;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    // a transform group no smaller than every job's count gives the same
    // work lists as the larger one before it
    auto same_lists = [&](int tgroup) {
      for (const Job &j : b.jobs)
        if (j.ntrans > std::max(1, batch_tmax(batch_kind(j), b.contr) >> tgroup)) return false;
      return true;
    };
    for (int tgroup = 0; tgroup <= 2; ++tgroup) {
      if (tgroup > 0 && same_lists(tgroup)) continue;
      // dispatch: persistent workgroups per CU 1, 2, 4; syntheses also flat
      for (int wpc : {1, 2, 4, 0})
        for (int shape = 0; shape <= 1; ++shape) {
          if (wpc == 0 && !b.synth) continue;
          b.tiles = BatchTiles{tgroup, wpc ? wpc : 2, shape, wpc == 0 ? 1 : 0};
          build_work(b, ncu);
          batch_run(c, b, b.cap);
          HIP_CHECK(hipEventRecord(e0, b.cap));
          batch_run(c, b, b.cap);
          HIP_CHECK(hipEventRecord(e1, b.cap));
          HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          if (ms < tbest) { tbest = ms; best = b.tiles; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_store(key, best);
  }
  b.tiles = best;
  build_work(b, ncu);
}

} // namespace shp

using namespace shp;

// Whether sorted intervals [first, second) cover [lo, hi] once, without gaps.
static bool covers(std::vector<std::pair<int64_t, int64_t>> &iv, int64_t lo, int64_t hi)
{
  if (iv.empty()) return false;
  std::sort(iv.begin(), iv.end());
  int64_t end = lo;
  for (const auto &x : iv) {
    if (x.first != end) return false;
    end = x.second;
  }
  return end == hi + 1;
}

// Every entry of the map range [lo, hi] (pixel units) belongs to one ring:
// then a map that is only written needs no upload before the download.
static bool map_dense(const sharp_geom_info *g, int64_t lo, int64_t hi)
{
  std::vector<std::pair<int64_t, int64_t>> iv;
  for (int i = 0; i < g->npairs; ++i)
    for (const sharp_ringinfo *r : {&g->pair[i].r1, &g->pair[i].r2}) {
      if (r->nph <= 0) continue;
      if (r->stride != 1) return false;
      iv.emplace_back(r->ofs, r->ofs + r->nph);
    }
  return covers(iv, lo, hi);
}

// Every entry of the a_lm range [lo, hi] (real units) is an a_lm element.
static bool alm_dense(const sharp_alm_info *a, int64_t lo, int64_t hi)
{
  const bool packed = a->flags & SHARP_PACKED;
  std::vector<std::pair<int64_t, int64_t>> iv;
  for (int mi = 0; mi < a->nm; ++mi) {
    const int m = a->mval[mi];
    if (m > a->lmax) continue;
    const bool realonly = packed && m == 0;
    int64_t ofs = a->mvstart[mi], stride = a->stride;
    if (!packed) ofs *= 2;
    if (!realonly) stride *= 2;
    if (stride != (realonly ? 1 : 2)) return false;
    iv.emplace_back(ofs + (int64_t)m * stride, ofs + (int64_t)(a->lmax + 1) * stride);
  }
  return covers(iv, lo, hi);
}

extern "C" {

void sharp_execute(sharp_jobtype type, int spin, void *alm, void *map,
                   const sharp_geom_info *geom_info, const sharp_alm_info *alm_info,
                   int ntrans, int flags, double *time, unsigned long long *opcnt)
{
  const auto t0 = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lock(g_mutex);
  const Job j = make_job(type, spin, geom_info, alm_info, ntrans, flags);
  Ctx &c = context();
  const int ncomp_alm = j.ntrans * j.nalm, ncomp_map = j.ntrans * j.nmaps;
  const size_t esz = (j.flags & SHARP_DP) ? sizeof(double) : sizeof(float);

  int64_t alo, ahi, mlo, mhi;
  alm_extent(alm_info, alo, ahi);
  map_extent(geom_info, mlo, mhi);
  const bool nofft = j.flags & SHARP_NO_FFT;
  const int64_t mreal_lo = nofft ? 2 * mlo : mlo, mreal_hi = nofft ? 2 * mhi + 1 : mhi;
  const size_t an = (size_t)(ahi - alo + 1), mn = (size_t)(mreal_hi - mreal_lo + 1);
  // (written from the copy stream too: plain allocation, not stream-ordered)
  char *da = static_cast<char *>(c.stage_alm.get(an * esz * ncomp_alm));
  char *dm = static_cast<char *>(c.stage_map.get(mn * esz * ncomp_map));
  void **halm = static_cast<void **>(alm), **hmap = static_cast<void **>(map);
  std::vector<void *> dalm(ncomp_alm), dmap(ncomp_map);
  hipStream_t st = c.stream;
  // the output array is uploaded only when the download could otherwise
  // overwrite host entries the transform does not write (SHARP_ADD, or gaps
  // in its range)
  const bool synth = j.type == SHARP_ALM2MAP, add = (j.flags & SHARP_ADD) != 0;
  const bool up_alm = synth || add || !alm_dense(alm_info, alo, ahi);
  const bool up_map = !synth || add || !map_dense(geom_info, mlo, mhi);
  for (int i = 0; i < ncomp_alm; ++i) dalm[i] = da + i * an * esz;
  for (int i = 0; i < ncomp_map; ++i) dmap[i] = dm + i * mn * esz;
  // the segments of transforms [t0, t1)
  auto segs = [&](int t0, int t1, std::vector<HostStage::Seg> &in,
                  std::vector<HostStage::Seg> &out) {
    for (int i = t0 * j.nalm; i < t1 * j.nalm; ++i) {
      char *h = static_cast<char *>(halm[i]) + alo * (int64_t)esz;
      if (up_alm) in.push_back({dalm[i], h, an * esz});
      if (!synth) out.push_back({h, dalm[i], an * esz});
    }
    for (int i = t0 * j.nmaps; i < t1 * j.nmaps; ++i) {
      char *h = static_cast<char *>(hmap[i]) + mreal_lo * (int64_t)esz;
      if (up_map) in.push_back({dmap[i], h, mn * esz});
      if (synth) out.push_back({h, dmap[i], mn * esz});
    }
  };
  // Inputs go through the pinned ring (host threads fill one chunk while the
  // DMA engine moves the previous one), then the transform, then the outputs
  // back the same way.  A large call runs as two groups of transforms, so
  // that the host copies and DMA of one group overlap the device work of the
  // other: upload A, transform A | upload B, transform B | download A,
  // download B.
  const size_t bytes = (size_t)(up_alm || !synth ? an * ncomp_alm : 0) * esz +
                       (size_t)(up_map || synth ? mn * ncomp_map : 0) * esz;
  static const int groups = [] {
    const char *e = std::getenv("SHARP_HIP_HOST_GROUPS");
    return e ? std::max(1, std::min(2, std::atoi(e))) : 2;
  }();
  const int ng = (groups > 1 && j.ntrans >= 2 && bytes >= (size_t(256) << 20)) ? 2 : 1;
  std::vector<HostStage::Seg> out_prev;
  for (int g = 0; g < ng; ++g) {
    const int t0 = j.ntrans * g / ng, t1 = j.ntrans * (g + 1) / ng;
    std::vector<HostStage::Seg> in, out;
    segs(t0, t1, in, out);
    Job jg = j;
    jg.ntrans = t1 - t0;
    c.host.upload(in, st);
    run_device(c, jg, dalm.data() + t0 * j.nalm, dmap.data() + t0 * j.nmaps, alo, mlo, st,
               nullptr);
    HIP_CHECK(hipEventRecord(c.host.done[g], st));
    if (g > 0) c.host.download_after(out_prev, c.host.done[g - 1]);
    out_prev.swap(out);
  }
  c.host.download_after(out_prev, c.host.done[ng - 1]);
  // libsharp's operation count of the whole call
  if (opcnt) {
    GeomPlan &gp = get_geom(c, j.g);
    *opcnt = op_count(j, gp, get_mlim(gp, j.lmax, j.spin));
  }
  if (time)
    *time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void sharp_hip_execute_device(sharp_jobtype type, int spin, void *alm, void *map,
                              const sharp_geom_info *geom_info,
                              const sharp_alm_info *alm_info, int ntrans, int flags,
                              void *stream, unsigned long long *opcnt)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  const Job j = make_job(type, spin, geom_info, alm_info, ntrans, flags);
  Ctx &c = context();
  // NULL is the device's null (default) stream, as everywhere in HIP; the
  // library's own stream is only used by the host-pointer sharp_execute
  hipStream_t st = static_cast<hipStream_t>(stream);
  run_device(c, j, static_cast<void *const *>(alm), static_cast<void *const *>(map),
             0, 0, st, opcnt);
}

void sharp_hip_execute_device_noalloc(sharp_jobtype type, int spin, void *alm, void *map,
                                      const sharp_geom_info *geom_info,
                                      const sharp_alm_info *alm_info, int ntrans, int flags,
                                      void *stream, unsigned long long *opcnt)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  const Job j = make_job(type, spin, geom_info, alm_info, ntrans, flags);
  // the call's work memory, laid out in the scratch area; sized from the job
  // alone and checked before any device call
  const int nslots = geom_info->npairs;
  const int gs = mix_group_slots(j.spin);
  const int nmb = (j.mmax + MB) / MB;
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t off = 0;
  const size_t o_phase = off;
  off += align((size_t)j.ntrans * j.nmaps * nmb * (size_t)nslots * 2 * MB * sizeof(double2));
  const size_t o_almtmp = off;
  off += align(almtmp_elems(j) * sizeof(double2));
  const size_t o_part = off;
  off += align(part_bytes(j, (nslots + gs - 1) / gs));
  const size_t o_bsplit = off, o_bsbase = off + (uses_bsplit(j) ? align((size_t)bs_words(j) * 4) : 0);
  if (uses_bsplit(j)) off = o_bsbase + align((size_t)(j.mmax + 1) * sizeof(int64_t));
  if (off > kNoallocScratchBytes) {
    char b[320];
    std::snprintf(b, sizeof(b),
                  "sharp_hip_execute_device_noalloc: this job needs %zu bytes of work memory, "
                  "more than the %zu of the fixed scratch area (use sharp_hip_execute_device)",
                  off, kNoallocScratchBytes);
    SHP_FAIL(b);
  }
  int dev = 0;
  char *base = nullptr;
  if (hipGetDevice(&dev) == hipSuccess && g_ctx.count(dev))
    base = static_cast<char *>(noalloc_scratch(dev));
  if (!base)
    SHP_FAIL("sharp_hip_execute_device_noalloc: call sharp_hip_prepare on this device first");
  Ctx &c = context();
  // long-ring FFT slots (the tables of the geometry must be prepared)
  g_noalloc = true;
  const GeomPlan &gp = get_geom(c, j.g);
  const int lslots = std::min(long_slots(gp, j), kNoallocLongSlots);
  const size_t o_long = off;
  off += align(long_scratch_bytes(gp, lslots));
  if (off > kNoallocScratchBytes) {
    char b[320];
    std::snprintf(b, sizeof(b),
                  "sharp_hip_execute_device_noalloc: this job needs %zu bytes of work memory, "
                  "more than the %zu of the fixed scratch area (use sharp_hip_execute_device)",
                  off, kNoallocScratchBytes);
    SHP_FAIL(b);
  }
  WorkPtrs w;
  w.lscratch = reinterpret_cast<double2 *>(base + o_long);
  w.lslots = lslots;
  w.phase = reinterpret_cast<double2 *>(base + o_phase);
  w.almtmp = reinterpret_cast<double2 *>(base + o_almtmp);
  w.part = base + o_part;
  w.bsplit = reinterpret_cast<unsigned *>(base + o_bsplit);
  w.bs_base = reinterpret_cast<int64_t *>(base + o_bsbase);
  // no per-stage events either: the profiler creates them on demand
  const bool prof = c.prof.on;
  c.prof.on = false;
  run_device(c, j, static_cast<void *const *>(alm), static_cast<void *const *>(map), 0, 0,
             static_cast<hipStream_t>(stream), opcnt, &w);
  g_noalloc = false;
  c.prof.on = prof;
}

void sharp_hip_alm2map_adjoint_device(int spin, void *alm, void *map,
                                      const sharp_geom_info *geom_info,
                                      const sharp_alm_info *alm_info, int ntrans,
                                      int flags, void *stream)
{
  sharp_hip_execute_device(SHARP_Yt, spin, alm, map, geom_info, alm_info, ntrans,
                           flags, stream, nullptr);
}

void sharp_hip_map2alm_adjoint_device(int spin, void *alm, void *map,
                                      const sharp_geom_info *geom_info,
                                      const sharp_alm_info *alm_info, int ntrans,
                                      int flags, void *stream)
{
  sharp_hip_execute_device(SHARP_WY, spin, alm, map, geom_info, alm_info, ntrans,
                           flags, stream, nullptr);
}

void sharp_hip_execute_scaled_device(sharp_jobtype type, int spin, void *alm, void *map,
                                     const sharp_geom_info *geom_info,
                                     const sharp_alm_info *alm_info, int ntrans, int flags,
                                     const double *lfac, void *stream)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  Job j = make_job(type, spin, geom_info, alm_info, ntrans, flags);
  if (!lfac) SHP_FAIL("per-l factors missing");
  j.lfac = lfac;
  Ctx &c = context();
  run_device(c, j, static_cast<void *const *>(alm), static_cast<void *const *>(map), 0, 0,
             static_cast<hipStream_t>(stream), nullptr);
}

void sharp_hip_map2alm_iter_device(int spin, void *alm, void *map,
                                   const sharp_geom_info *geom_info,
                                   const sharp_alm_info *alm_info, int ntrans, int flags,
                                   int num_iter, void *stream)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (num_iter < 0) SHP_FAIL("bad number of iterations");
  if (flags & SHARP_ADD) SHP_FAIL("SHARP_ADD is not supported by the iterated analysis");
  const Job ja = make_job(SHARP_MAP2ALM, spin, geom_info, alm_info, ntrans, flags);
  Ctx &c = context();
  hipStream_t st = static_cast<hipStream_t>(stream);
  void *const *dalm = static_cast<void *const *>(alm);
  void *const *dmap = static_cast<void *const *>(map);
  run_device(c, ja, dalm, dmap, 0, 0, st, nullptr);
  if (num_iter == 0) return;
  const Job js = make_job(SHARP_ALM2MAP, spin, geom_info, alm_info, ntrans, flags);
  const Job jadd = make_job(SHARP_MAP2ALM, spin, geom_info, alm_info, ntrans, flags | SHARP_ADD);
  // residual maps over the pixel extent of the geometry (reals; complex
  // pixels under SHARP_NO_FFT)
  int64_t mlo, mhi;
  map_extent(geom_info, mlo, mhi);
  const bool nofft = flags & SHARP_NO_FFT;
  const int64_t rlo = nofft ? 2 * mlo : mlo, rhi = nofft ? 2 * mhi + 1 : mhi;
  const int64_t n = rhi - rlo + 1;
  const int ncomp = ja.ntrans * ja.nmaps;
  const size_t esz = (flags & SHARP_DP) ? sizeof(double) : sizeof(float);
  char *rb = static_cast<char *>(c.bufs(st).resid.get((size_t)n * esz * ncomp, st, true));
  std::vector<void *> rmap(ncomp);
  PtrTab src{}, dst{};
  for (int i = 0; i < ncomp; ++i) {
    rmap[i] = rb + (size_t)i * n * esz;
    src.p[i] = static_cast<char *>(dmap[i]) + rlo * (int64_t)esz;
    dst.p[i] = rmap[i];
  }
  for (int it = 0; it < num_iter; ++it) {
    // Healpix_cxx/alm_healpix_tools.cc:77-84: map2 = map - alm2map(alm);
    // alm += map2alm(map2)
    run_device(c, js, dalm, rmap.data(), 0, mlo, st, nullptr);
    timed(c, st, "map_residual", 0., 3. * (double)n * esz * ncomp,
          [&] { launch_map_residual(src, dst, ncomp, n, (flags & SHARP_DP) != 0, st); });
    run_device(c, jadd, dalm, rmap.data(), 0, mlo, st, nullptr);
  }
}

sharp_hip_batch *sharp_hip_batch_create(int njobs, const sharp_hip_job *jobs, int flags,
                                        int use_graph)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (njobs <= 0 || !jobs) SHP_FAIL("empty batch");
  Ctx &c = context();
  auto b = std::make_unique<BatchPlan>();
  b->device = c.device;
  b->contr = is_mixed(flags) ? mixed_contr(flags) : kContrF64;
  b->use_graph = use_graph != 0;
  for (int i = 0; i < njobs; ++i) {
    const sharp_hip_job &q = jobs[i];
    const Job j = make_job(q.type, q.spin, q.geom_info, q.alm_info, q.ntrans, flags);
    const bool syn = (j.type == SHARP_ALM2MAP);
    if (i == 0) b->synth = syn;
    else if (syn != b->synth) SHP_FAIL("batch jobs must all be syntheses or all analyses");
    b->jobs.push_back(j);
    void *const *ap = static_cast<void *const *>(q.alm);
    void *const *mp = static_cast<void *const *>(q.map);
    b->almp.emplace_back(ap, ap + (size_t)j.ntrans * j.nalm);
    b->mapp.emplace_back(mp, mp + (size_t)j.ntrans * j.nmaps);
  }
  const size_t n = b->jobs.size();
  for (auto *v : {&b->phase, &b->almtmp, &b->part, &b->bsplit, &b->gmax2}) v->resize(n);
  for (int wi = 0; wi < 2; ++wi) {
    b->ngroups[wi].resize(n);
    b->dgmax[wi].resize(n);
  }
  for (size_t i = 0; i < n; ++i) {
    const Job &j = b->jobs[i];
    GeomPlan &gp = get_geom(c, j.g);
    double2 *ph = static_cast<double2 *>(b->phase[i].get(phase_elems(j, gp) * sizeof(double2)));
    double2 *at = static_cast<double2 *>(b->almtmp[i].get(almtmp_elems(j) * sizeof(double2)));
    b->st.push_back(setup_job(c, j, b->almp[i].data(), b->mapp[i].data(), 0, 0, ph, at, nullptr,
                              c.stream));
    JobState &s = b->st.back();
    // ring groups of the two analysis ring blocks (4 and 2 wavefronts) and
    // the largest m limit of each
    b->ngroups[0][i] = (int)s.ml->hgmax.size();
    b->dgmax[0][i] = s.ml->gmax;
    const int gs2 = 128, ng2 = (gp.nslots + gs2 - 1) / gs2;
    std::vector<int> hg2(ng2, -1);
    for (int k = 0; k < gp.nslots; ++k) hg2[k / gs2] = std::max(hg2[k / gs2], s.ml->h[k]);
    int *d2 = static_cast<int *>(b->gmax2[i].get(hg2.size() * sizeof(int)));
    HIP_CHECK(hipMemcpy(d2, hg2.data(), hg2.size() * sizeof(int), hipMemcpyHostToDevice));
    b->ngroups[1][i] = ng2;
    b->dgmax[1][i] = d2;
    if (!b->synth) {
      // partial sums of every transform of the job, [group][row][2*ncomp_alm],
      // with room for the groups of the smaller ring block
      const int64_t nrows = tri_ofs(j.mmax + 1, j.lmax);
      const size_t esz = b->contr == kContrF64 ? sizeof(double) : sizeof(float);
      s.la.part = b->part[i].get((size_t)std::max(ng2, b->ngroups[0][i]) * (size_t)nrows * 2 *
                                 s.ncomp_alm * esz);
      s.la.part_ld = 2 * s.ncomp_alm;
      s.aa.part = s.la.part;
      s.aa.part_f64 = b->contr == kContrF64 ? 1 : 0;
      s.aa.part_ld = s.la.part_ld;
      s.aa.part_c0 = 0;
      s.aa.c0 = 0;
      s.aa.npc = s.ncomp_alm;
    }
    if (uses_bsplit(j)) {
      // the BF16 spin-s synthesis reads the a_lm split into BF16 pieces by
      // the gather
      const int ncol = bs_ncol(j);
      BsPlan &bp = get_bsplan(c, j.lmax, j.mmax, j.spin, ncol);
      s.keep.push_back(owner(c.bsplans, [&](const BsPlan &x) { return &x == &bp; }));
      unsigned *bs = static_cast<unsigned *>(b->bsplit[i].get((size_t)bp.total * sizeof(unsigned)));
      s.aa.bsplit = bs;
      s.la.bsplit = bs;
      s.aa.bs_base = s.la.bs_base = bp.d;
      s.aa.bs_ncol = s.la.bs_ncol = ncol;
    }
  }
  // long-ring scratch: the largest any job needs, shared in turn
  size_t lbytes = 0;
  for (auto &s : b->st) {
    s.lslots = long_slots(*s.gp, s.j);
    lbytes = std::max(lbytes, long_scratch_bytes(*s.gp, s.lslots));
  }
  if (lbytes) {
    double2 *ls = static_cast<double2 *>(b->lscratch.get(lbytes));
    for (auto &s : b->st) s.lscratch = ls;
  }
  for (int wi = 0; wi < 2; ++wi) {
    std::vector<LegArgs> las;
    for (size_t i = 0; i < n; ++i) {
      LegArgs la = b->st[i].la;
      la.gmax = b->dgmax[wi][i];
      la.ngroups = b->ngroups[wi][i];
      las.push_back(la);
    }
    b->djobs[wi] = upload(las);
  }
  HIP_CHECK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  int ncu = 256;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess)
    ncu = 256;
  const bool prof = c.prof.on;
  c.prof.on = false;
  batch_tune(c, *b, ncu);
  batch_run(c, *b, b->cap);   // settles every lazily built plan before capture
  HIP_CHECK(hipStreamSynchronize(b->cap));
  if (b->use_graph) {
    HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    batch_run(c, *b, b->cap);
    HIP_CHECK(hipStreamEndCapture(b->cap, &b->graph));
    HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  c.prof.on = prof;
  return reinterpret_cast<sharp_hip_batch *>(b.release());
}

void sharp_hip_batch_execute(sharp_hip_batch *batch, void *stream)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  BatchPlan *b = reinterpret_cast<BatchPlan *>(batch);
  if (!b) SHP_FAIL("null batch");
  Ctx &c = context();
  if (c.device != b->device) SHP_FAIL("batch executed on another device than it was created on");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->exec) {
    timed(c, st, "batch_graph", 0., 0., [&] { HIP_CHECK(hipGraphLaunch(b->exec, st)); });
  } else {
    batch_run(c, *b, st);
  }
}

void sharp_hip_batch_config(const sharp_hip_batch *batch, int *tgroup, int *wg_per_cu,
                            int *graph)
{
  const BatchPlan *b = reinterpret_cast<const BatchPlan *>(batch);
  if (!b) SHP_FAIL("null batch");
  if (tgroup)
    *tgroup = b->jobs.empty() ? 0
                              : std::max(1, batch_tmax(batch_kind(b->jobs[0]), b->contr) >>
                                                b->tiles.tgroup);
  if (wg_per_cu) *wg_per_cu = (b->synth && b->tiles.flat) ? 0 : b->tiles.wg_per_cu;
  if (graph) *graph = b->exec ? 1 : 0;
}

void sharp_hip_batch_tiles(const sharp_hip_batch *batch, int *mblock, int *ring_waves)
{
  const BatchPlan *b = reinterpret_cast<const BatchPlan *>(batch);
  if (!b) SHP_FAIL("null batch");
  if (mblock) *mblock = b->synth ? 2 * batch_waves(b->tiles.shape) : 0;
  if (ring_waves) *ring_waves = b->synth ? 0 : batch_waves(b->tiles.shape);
}

void sharp_hip_batch_destroy(sharp_hip_batch *batch)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  BatchPlan *b = reinterpret_cast<BatchPlan *>(batch);
  if (!b) return;
  (void)hipStreamSynchronize(b->cap);
  HIP_CHECK(hipDeviceSynchronize());
  delete b;
}

void sharp_hip_prepare(const sharp_geom_info *geom_info,
                       const sharp_alm_info *alm_info, int spin)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  Ctx &c = context();
  const int mmax = alm_info->nm - 1;
  GeomPlan &gp = get_geom(c, geom_info);
  get_mlim(gp, alm_info->lmax, spin);
  // libsharp's table (FP64 form) and the unit-coefficient one (mixed forms)
  ensure_unit_table(get_ylm(c, alm_info->lmax, mmax, spin));
  get_norm(c, alm_info->lmax, spin, 0);
  if (spin == 1) get_norm(c, alm_info->lmax, spin, 1);   // ALM2MAP_DERIV1
  get_almplan(c, alm_info, mmax);
  // the address of this device's scratch area (sharp_hip_execute_device_noalloc)
  noalloc_setup();
}

void sharp_hip_synchronize(void *stream)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  Ctx &c = context();
  (void)c;
  HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

void sharp_hip_release(void)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  auto it = g_ctx.find(dev);
  if (it == g_ctx.end()) return;
  Ctx &c = *it->second;
  (void)hipDeviceSynchronize();
  c.geoms.clear();
  c.ylms.clear();
  c.norms.clear();
  c.alms.clear();
  c.bsplans.clear();
  for (auto &kv : c.sbufs) {
    Ctx::StreamBufs &b = *kv.second;
    for (DevBuf *d : {&b.phase, &b.almtmp, &b.part, &b.resid, &b.lfac, &b.bsplit, &b.lscratch})
      d->release();
  }
  c.sbufs.clear();
  c.stage_alm.release();
  c.stage_map.release();
  c.host.release();
}

void sharp_hip_profile(int enable)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  Ctx &c = context();
  c.prof.drain();
  c.prof.stats.clear();
  c.prof.on = enable != 0;
}

int sharp_hip_profile_stats(int maxn, char *names, double *ms, long long *launches,
                            double *flops, double *bytes)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  Ctx &c = context();
  c.prof.drain();
  const int n = (int)c.prof.stats.size();
  for (int i = 0; i < n && i < maxn; ++i) {
    const auto &s = c.prof.stats[i];
    std::snprintf(names + 32 * i, 32, "%s", s.name.c_str());
    ms[i] = s.ms;
    launches[i] = s.launches;
    flops[i] = s.flops;
    bytes[i] = s.bytes;
  }
  return n;
}

int sharp_hip_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

} // extern "C"

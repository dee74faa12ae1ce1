// ring_fft.hip -- batched ring FFTs for the ragged rings of a map.
//
// One workgroup transforms one ring (for one or several components).  The
// ring's data never leaves LDS between loading and storing:
//
//   synthesis (phase -> ring), libsharp/sharp.c:253-303:
//     1. fold   c_m = w e^{i m phi0} phase_m  (m = 0..mmax) onto the
//               Hermitian bins C_k (k = m mod n, conjugated for -m): the
//               aliasing fold of rings with n < 2 mmax + 1;
//     2. pack   even n: the half-length complex sequence whose inverse FFT
//               interleaves even and odd pixels; odd n: the full Hermitian
//               sequence;
//     3. FFT    inverse complex FFT of length N (= n/2 or n) in LDS;
//     4. store  pixels (added to the map with SHARP_ADD).
//   analysis (ring -> phase), libsharp/sharp.c:305-353: the reverse, ending
//     with phase_m = w e^{-i m phi0} X_{m mod n} (conjugated past n/2).
//
// Complex FFTs of length N are Stockham autosort passes of radix 8, 4, 2, 3,
// 5 and 7, in place in one LDS buffer (every pass reads all its inputs into
// registers, barriers, then writes).  A length with another prime factor is
// done by Bluestein's algorithm: a circular convolution of length
// M >= 2N-1 (a power of two) with the chirp e^{-i pi j^2 / N}; the chirp and
// the FFT of the convolution kernel are precomputed once per (N, M) into a
// table in HBM, and the pointwise products ride on the first and last
// passes of the two length-M transforms.  The transform helpers are forced
// inline: as device calls they cost registers, spills and a fifth of the
// kernels' time.  Twiddles of a pass come from one sincos per butterfly; the
// azimuthal rotations e^{+-i m phi0} from one sincos per group of bins,
// advanced by complex multiplication.
//
// Rings are grouped by the LDS buffer they need (512 ... 8192 complex
// doubles); each group is one launch of a kernel instantiated for that size
// with S/8 threads, so every thread carries at most 8 complex values per pass.

#include <hip/hip_runtime.h>
#include "kernels.hpp"
#include "sharp_lowlevel.h"
#include "devmath.hpp"

#include <algorithm>
#include <map>

namespace shp {

namespace {

constexpr double kSqrtHalf = 0.707106781186547572737310929369;
constexpr double kSqrt2 = 1.414213562373095145474621858739;

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 conj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// The FFT buffer in LDS with a 16-byte pad after every 8 entries: entry i at
// i + i/8.  The first Stockham pass writes the outputs of a butterfly to
// consecutive entries, R entries apart from lane to lane; unpadded, the lanes
// of a radix-8 pass then land on the same banks.
__host__ __device__ constexpr int lds_pad(int n) { return n + (n >> 3); }
struct Lds {
  double2 *p;
  __device__ double2 &operator[](int i) const { return p[i + (i >> 3)]; }
};

// multiply by sign*i
__device__ inline double2 mul_si(double2 a, int sign)
{
  return sign > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
// e^{i pi x} for |x| <= 2
__device__ inline double2 expipi(double x)
{
  return expi_small(3.141592653589793 * x);
}

// a wave-uniform value in scalar registers (the compiler keeps FP64 values
// computed by vector instructions in vector registers otherwise)
__device__ inline double uniform(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double2 uniform(double2 v) { return make_double2(uniform(v.x), uniform(v.y)); }

// The roots of unity e^{2 pi i j / kTwN}, written once per device
// (ensure_twiddles): the twiddles of the passes with compile-time strides
// (every Ns*R divides kTwN) and of the even-ring packing of the
// power-of-two classes (n divides kTwN) are table reads instead of sincos.
constexpr int kTwN = 16384;
__device__ double2 g_twiddle[kTwN];

template <typename T> __device__ inline double ld(const void *p, int64_t i)
{
  return (double)static_cast<const T *>(p)[i];
}
template <typename T> __device__ inline void st(void *p, int64_t i, double v)
{
  static_cast<T *>(p)[i] = (T)v;
}
// elements i, i+1 as one vector access: needs unit stride and p + i aligned
// to two elements for every even i (checked once per ring by pair_aligned)
template <typename T> __device__ inline bool pair_aligned(const void *p, int64_t ofs,
                                                          int64_t stride)
{
  return stride == 1 && (ofs & 1) == 0 &&
         (reinterpret_cast<uintptr_t>(p) & (2 * sizeof(T) - 1)) == 0;
}
template <typename T> __device__ inline double2 ld_pair(const void *p, int64_t i)
{
  if constexpr (sizeof(T) == 8) {
    return *reinterpret_cast<const double2 *>(static_cast<const double *>(p) + i);
  } else {
    const float2 v = *reinterpret_cast<const float2 *>(static_cast<const float *>(p) + i);
    return make_double2(v.x, v.y);
  }
}
template <typename T> __device__ inline void st_pair(void *p, int64_t i, double2 v)
{
  if constexpr (sizeof(T) == 8) {
    *reinterpret_cast<double2 *>(static_cast<double *>(p) + i) = v;
  } else {
    *reinterpret_cast<float2 *>(static_cast<float *>(p) + i) = make_float2((float)v.x, (float)v.y);
  }
}

// --- small DFTs: X_q = sum_r v_r e^{sign 2 pi i r q / R} ---------------------
template <int R> struct Dft;

template <> struct Dft<2> {
  static __device__ inline void run(double2 *v, int)
  {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <> struct Dft<4> {
  static __device__ inline void run(double2 *v, int sign)
  {
    const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const double2 t2 = cadd(v[1], v[3]), t3 = mul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <> struct Dft<8> {
  static __device__ inline void run(double2 *v, int sign)
  {
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4>::run(e, sign);
    Dft<4>::run(o, sign);
    // o_k *= e^{sign 2 pi i k / 8}
    const double h = kSqrtHalf;
    o[1] = cmul(o[1], make_double2(h, sign * h));
    o[2] = mul_si(o[2], sign);
    o[3] = cmul(o[3], make_double2(-h, sign * h));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

template <> struct Dft<16> {
  static __device__ inline void run(double2 *v, int sign)
  {
    double2 e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { e[k] = v[2 * k]; o[k] = v[2 * k + 1]; }
    Dft<8>::run(e, sign);
    Dft<8>::run(o, sign);
    // o_k *= e^{sign 2 pi i k / 16}: cos and sin of pi/8
    const double c = 0.923879532511286756128, s = 0.382683432365089771728, h = kSqrtHalf;
    o[1] = cmul(o[1], make_double2(c, sign * s));
    o[2] = cmul(o[2], make_double2(h, sign * h));
    o[3] = cmul(o[3], make_double2(s, sign * c));
    o[4] = mul_si(o[4], sign);
    o[5] = cmul(o[5], make_double2(-s, sign * c));
    o[6] = cmul(o[6], make_double2(-h, sign * h));
    o[7] = cmul(o[7], make_double2(-c, sign * s));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 8] = csub(e[k], o[k]);
    }
  }
};

template <> struct Dft<3> {
  static __device__ inline void run(double2 *v, int sign)
  {
    const double s3 = 0.866025403784438646763723170752936183;  // sqrt(3)/2
    const double2 a = v[0], b = v[1], c = v[2];
    const double2 bc = cadd(b, c);
    const double2 t = make_double2(a.x - 0.5 * bc.x, a.y - 0.5 * bc.y);
    const double2 u = mul_si(cscale(csub(b, c), s3), sign);
    v[0] = cadd(a, bc);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  }
};

// odd prime radix R = 5, 7 by the symmetric pairing of r and R-r:
//   X_q, X_{R-q} = v_0 + sum_r (v_r + v_{R-r}) cos(2 pi r q / R)
//                  +- sign i sum_r (v_r - v_{R-r}) sin(2 pi r q / R)
struct Tw5 {
  static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475,
                                  -0.8090169943749475, 0.30901699437494745};
  static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924731,
                                  -0.5877852522924731, -0.9510565162951535};
};
struct Tw7 {
  static constexpr double c[7] = {1.0, 0.6234898018587335, -0.2225209339563144,
                                  -0.9009688679024191, -0.9009688679024191,
                                  -0.2225209339563144, 0.6234898018587335};
  static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236,
                                  0.4338837391175581, -0.4338837391175581,
                                  -0.9749279121818236, -0.7818314824680298};
};
template <int R, typename TW> struct DftOdd {
  static __device__ inline void run(double2 *v, int sign)
  {
    constexpr int H = (R - 1) / 2;
    double2 a[H + 1], b[H + 1];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      a[r] = cadd(v[r], v[R - r]);
      b[r] = csub(v[r], v[R - r]);
    }
    double2 x0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; ++r) x0 = cadd(x0, a[r]);
    double2 out[R];
    out[0] = x0;
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      double2 A = v[0], B = make_double2(0., 0.);
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const int k = (r * q) % R;
        A.x = fma(a[r].x, TW::c[k], A.x);
        A.y = fma(a[r].y, TW::c[k], A.y);
        B.x = fma(b[r].x, TW::s[k], B.x);
        B.y = fma(b[r].y, TW::s[k], B.y);
      }
      const double2 iB = mul_si(B, sign);
      out[q] = cadd(A, iB);
      out[R - q] = csub(A, iB);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
  }
};
template <> struct Dft<5> : DftOdd<5, Tw5> {};
template <> struct Dft<7> : DftOdd<7, Tw7> {};

// --- one Stockham pass of radix R over buf[0..N) -----------------------------
// With FIRST the pass reads its inputs through `pre` (pre(buf, i) instead of
// buf[i]); with LAST it passes its outputs through `post` (post(i, v)
// updates v and returns whether buf[i] is written).  The Bluestein
// convolution folds its pointwise products into them.
struct PlainLoad {
  __device__ double2 operator()(Lds b, int i) const { return b[i]; }
};
struct PlainStore {
  __device__ bool operator()(int, double2 &) const { return true; }
};

// TWT: Ns*R divides kTwN and the twiddles come from g_twiddle.  HIN: the
// inputs r >= R/2 of every butterfly are zero (the first pass of a
// transform whose upper half is zero padding); HOUT: the outputs r >= R/2
// are not wanted (the last pass of a transform of which only the lower half
// is kept).
template <int NT, int E, int R, bool FIRST = false, bool LAST = false,
          typename Pre = PlainLoad, typename Post = PlainStore, bool TWT = false,
          bool HIN = false, bool HOUT = false>
__device__ __forceinline__ void stockham_pass(Lds buf, int N, int Ns, int sign,
                                     const Pre &pre = Pre(), const Post &post = Post())
{
  constexpr int BPT = (E + R - 1) / R;
  // the thread index through an empty asm, so that nothing of this pass
  // (addresses, twiddles) is computed ahead of it: with compile-time
  // lengths and strides it all depends on the thread index alone, and the
  // compiler would hoist every pass's share out of the component loop and
  // spill it
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int nb = N / R;
  double2 v[BPT][R];
  // table twiddles are read ahead of the barrier (their latency hides
  // behind the LDS reads and the wait)
  double2 wt[TWT ? BPT : 1];
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int j = tid + i * NT;
    if (j < nb) {
      if constexpr (TWT) {
        if (Ns > 1) wt[i] = g_twiddle[(j & (Ns - 1)) * (kTwN / (Ns * R))];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (HIN && 2 * r >= R) v[i][r] = make_double2(0., 0.);
        else if constexpr (FIRST) v[i][r] = pre(buf, j + r * nb);
        else v[i][r] = buf[j + r * nb];
      }
    }
  }
  __syncthreads();
  // twiddle angle per unit of k, in units of pi; k = j mod Ns by a mask when
  // Ns is a power of two (every pass of radix 8, 4, 2), else from the FP32
  // quotient (j + 1/2) / Ns, whose fraction stays >= 1/(2 Ns) from an
  // integer while its rounding error is below 2^-22 j / Ns (j < 2^16)
  const double tscale = (double)sign * 2.0 / (double)(Ns * R);
  const bool pow2 = (Ns & (Ns - 1)) == 0;
  const float rns = 1.0f / (float)Ns;
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int j = tid + i * NT;
    if (j < nb) {
      const int k = pow2 ? (j & (Ns - 1)) : (j - Ns * (int)(((float)j + 0.5f) * rns));
      if (Ns > 1) {
        double2 w;
        if constexpr (TWT) w = sign > 0 ? wt[i] : conj(wt[i]);
        else w = expipi((double)k * tscale);
        double2 wr = w;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[i][r] = cmul(v[i][r], wr);
          if (r + 1 < R) wr = cmul(wr, w);
        }
      }
      Dft<R>::run(v[i], sign);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < (HOUT ? R / 2 : R); ++r) {
        if constexpr (LAST) {
          if (post(base + r * Ns, v[i][r])) buf[base + r * Ns] = v[i][r];
        } else {
          buf[base + r * Ns] = v[i][r];
        }
      }
    }
  }
  __syncthreads();
}

// complex FFT of length N (product of the plan radices) in place in LDS.
// The plan lists the radices grouped as 8..8, then 4 or 2, then 3s, 5s, 7s;
// the passes are laid out in that order as straight-line code (a switch
// inside one loop makes the register allocator keep every pass's values
// live at once).
// P2 (radix set of the plan): 0 any of 8, 4, 2, 3, 5, 7; 1 only 8, 4, 2 and
// (16 values per thread) 16: a power-of-two length; 2 only 8.  The passes a plan cannot have are left
// out of the kernel, which keeps its register count (and so its occupancy)
// at what the passes it runs need.
// Passes rad[s0 .. s1) starting from stride Ns; returns the stride after them.
template <int NT, int E, int P2>
__device__ __forceinline__ int lds_fft(Lds buf, int N, const int8_t *rad, int nrad, int sign, int s0 = 0,
                       int Ns = 1, int s1 = -1)
{
  if (s1 < 0) s1 = nrad;
  int s = s0;
  // radix 16 (plans of 16 values per thread only, ahead of the 8s)
  if constexpr (P2 < 2 && E >= 16)
    for (; s < s1 && rad[s] == 16; ++s, Ns *= 16) stockham_pass<NT, E, 16>(buf, N, Ns, sign);
  for (; s < s1 && rad[s] == 8; ++s, Ns *= 8) stockham_pass<NT, E, 8>(buf, N, Ns, sign);
  if constexpr (P2 < 2) {
    if (s < s1 && rad[s] == 4) { stockham_pass<NT, E, 4>(buf, N, Ns, sign); Ns *= 4; ++s; }
    if (s < s1 && rad[s] == 2) { stockham_pass<NT, E, 2>(buf, N, Ns, sign); Ns *= 2; ++s; }
  }
  if constexpr (P2 == 0) {
    for (; s < s1 && rad[s] == 3; ++s, Ns *= 3) stockham_pass<NT, E, 3>(buf, N, Ns, sign);
    for (; s < s1 && rad[s] == 5; ++s, Ns *= 5) stockham_pass<NT, E, 5>(buf, N, Ns, sign);
    for (; s < s1 && rad[s] == 7; ++s, Ns *= 7) stockham_pass<NT, E, 7>(buf, N, Ns, sign);
  }
  return Ns;
}

// FFT of a Bluestein length M (a power of two >= 64) whose plan lists the
// radices as 8, (8s, one 4 or 2), 8: the first and the last pass are of
// radix 8, with `pre` and `post` applied.
template <int NT, int E, int P2, typename Pre = PlainLoad, typename Post = PlainStore>
__device__ __forceinline__ void bluestein_fft(Lds buf, const RingFftPlan &P, int sign,
                              const Pre &pre = Pre(), const Post &post = Post())
{
  const int M = P.M;
  stockham_pass<NT, E, 8, true, false>(buf, M, 1, sign, pre);
  const int Ns = lds_fft<NT, E, P2>(buf, M, P.rad, P.nrad, sign, 1, 8, P.nrad - 1);
  stockham_pass<NT, E, 8, false, true>(buf, M, Ns, sign, PlainLoad(), post);
}

// Bluestein chirp h_j = e^{sign i pi j^2 / N} with the exact angle (table
// construction only; the transforms read it from the table)
__device__ inline double2 chirp(int64_t j, int N, int sign)
{
  const int64_t q = (j * j) % (2 * (int64_t)N);
  return expipi(sign * (double)q / (double)N);
}

// FFT of the convolution kernel b_d = e^{-sign i pi d^2 / N}, |d| < N,
// wrapped to length M, left in buf[0..M)
template <int NT, int E>
__device__ void bluestein_kernel(Lds buf, const RingFftPlan &P, int sign)
{
  const int N = P.N, M = P.M;
  for (int j = threadIdx.x; j < M; j += NT) {
    int d = -1;
    if (j < N) d = j;
    else if (j > M - N) d = M - j;
    buf[j] = (d >= 0) ? chirp(d, N, -sign) : make_double2(0., 0.);
  }
  __syncthreads();
  bluestein_fft<NT, E, 1>(buf, P, -1);
}

// The passes of a transform whose length is the class size S, known at
// compile time: in the classes of 1024 and more entries, every plan of a
// power-of-two radix set has that length (a power-of-two N in (S/2, S] is S;
// a Bluestein M is a power of two above the next smaller class), and its
// radices are those fft_plan lists (engine.cpp), which the host checks
// before the launch (fixed_fft_plan_ok).  With the lengths and strides
// constant, the index arithmetic and the range checks of the passes fold
// away.
template <int S, bool BLUE> struct FixedPasses { static constexpr int n = 0, r[5] = {}; };
template <> struct FixedPasses<8192, true> { static constexpr int n = 4, r[5] = {8, 16, 8, 8}; };
template <> struct FixedPasses<4096, true> { static constexpr int n = 4, r[5] = {8, 8, 8, 8}; };
template <> struct FixedPasses<2048, true> { static constexpr int n = 4, r[5] = {8, 8, 4, 8}; };
template <> struct FixedPasses<1024, true> { static constexpr int n = 4, r[5] = {8, 8, 2, 8}; };
template <> struct FixedPasses<8192, false> { static constexpr int n = 5, r[5] = {8, 8, 8, 8, 2}; };
template <> struct FixedPasses<4096, false> { static constexpr int n = 4, r[5] = {8, 8, 8, 8}; };
template <> struct FixedPasses<2048, false> { static constexpr int n = 4, r[5] = {8, 8, 8, 4}; };
template <> struct FixedPasses<1024, false> { static constexpr int n = 4, r[5] = {8, 8, 8, 2}; };
template <int S, int P2> constexpr bool kFixedLen = S >= 1024 && P2 >= 1;

// e^{2 pi i k / n} of the even-ring packing and unpacking: a table read in
// the power-of-two classes, where n = 2N = 2S
template <int S, int P2, bool BLUE>
__device__ inline double2 pack_twiddle(int k, int n)
{
  if constexpr (kFixedLen<S, P2> && !BLUE) {
    static_assert(kTwN % (2 * S) == 0, "twiddle table too short");
    return g_twiddle[k * (kTwN / (2 * S))];
  } else {
    return expipi((double)k * (2.0 / n));
  }
}

// passes I.. of the fixed plan, starting at stride NS; Bluestein transforms
// apply `pre` in their first pass and `post` in their last.  HIN: the input's
// upper half is zero; HOUT: only the output's lower half is wanted (a
// first pass of stride 1 reads input j + r S/R of butterfly j, and a last
// pass of stride S/R writes output j + r S/R: in both, r >= R/2 is the upper
// half).
template <int NT, int E, int S, bool BLUE, int I, int NS, bool HIN = false, bool HOUT = false,
          typename Pre, typename Post>
__device__ __forceinline__ void fixed_passes(Lds buf, int sign, const Pre &pre, const Post &post)
{
  typedef FixedPasses<S, BLUE> FP;
  if constexpr (I < FP::n) {
    constexpr int R = FP::r[I];
    constexpr bool first = I == 0, last = I == FP::n - 1;
    static_assert(kTwN % (NS * R) == 0, "twiddle table too short");
    static_assert(!(HIN && first) || NS == 1, "zero upper half needs the stride-1 first pass");
    static_assert(!(HOUT && last) || NS * R == S, "lower half needs the last pass");
    stockham_pass<NT, E, R, BLUE && first, BLUE && last, Pre, Post, true, HIN && first,
                  HOUT && last>(buf, S, NS, sign, pre, post);
    fixed_passes<NT, E, S, BLUE, I + 1, NS * R, HIN, HOUT>(buf, sign, pre, post);
  }
}

// length-N transform of buf[0..N) with the sign `sign`, smooth or Bluestein;
// Bt is the plan's block of the Bluestein table: the spectrum of the
// convolution kernel (entries 0..M/2, see bluestein_kernel), then the chirp
// e^{-i pi j^2/N} (entries 0..N/2)
template <int NT, int E, int S, bool BLUE, int P2>
__device__ __forceinline__ void ring_dft(Lds buf, const RingFftPlan &P, int sign, const double2 *Bt)
{
  constexpr bool FIX = kFixedLen<S, P2>;
  if constexpr (!BLUE) {
    if constexpr (FIX) fixed_passes<NT, E, S, false, 0, 1>(buf, sign, PlainLoad(), PlainStore());
    else lds_fft<NT, E, P2>(buf, P.N, P.rad, P.nrad, sign);
  } else {
  const int N = P.N, M = FIX ? S : P.M;
  // the table holds half of each symmetric sequence: the kernel spectrum
  // B_k = B_{M-k} for k <= M/2, the chirp c_j = (-1)^N c_{N-j} for j <= N/2
  const double2 *Ct = Bt + M / 2 + 1;
  auto ch = [&](int j) {
    const bool hi = 2 * j > N;
    double2 c = Ct[hi ? N - j : j];
    if (hi && (N & 1)) c = make_double2(-c.x, -c.y);
    return sign > 0 ? conj(c) : c;
  };
  // forward transform of the chirp-weighted input, zero-padded to M (read
  // in its first pass), times the kernel spectrum (in its last pass).  The
  // kernel of sign s is the conjugate of the kernel of sign -s with the
  // input conjugated: FFT(conj(x))_k = conj(FFT(x)_{-k}); the table holds
  // the spectrum for the analysis sign -1, and as B_{-k} = B_k the other
  // sign's factor is conj(B_k)
  auto chirped = [&](Lds b, int j) {
    return j < N ? cmul(b[j], ch(j)) : make_double2(0., 0.);
  };
  auto kernel = [&](int j, double2 &v) {
    const double2 b = Bt[2 * j > M ? M - j : j];
    v = cmul(v, sign > 0 ? conj(b) : b);
    return true;
  };
  // (a Bluestein length is at least 2N: the chirped input fills at most the
  // lower half, and only the lower half of the inverse is kept)
  if constexpr (FIX) fixed_passes<NT, E, S, true, 0, 1, true, false>(buf, -1, chirped, kernel);
  else bluestein_fft<NT, E, P2>(buf, P, -1, chirped, kernel);
  // inverse transform; only outputs j < N are kept, times chirp and 1/M
  const double inv = 1.0 / M;
  auto out = [&](int j, double2 &v) {
    if (j >= N) return false;
    v = cscale(cmul(v, ch(j)), inv);
    return true;
  };
  if constexpr (FIX) fixed_passes<NT, E, S, true, 0, 1, false, true>(buf, +1, PlainLoad(), out);
  else bluestein_fft<NT, E, P2>(buf, P, +1, PlainLoad(), out);
  }
}

// reduce a complex value over the G consecutive lanes that share a bin
__device__ inline double2 group_sum(double2 v, int G)
{
  for (int d = 1; d < G; d <<= 1) {
    v.x += __shfl_xor(v.x, d);
    v.y += __shfl_xor(v.y, d);
  }
  return v;
}

// Workgroup b transforms plan pair_plan(b): plans 2q and 2q+1 (the northern
// and southern ring of a pair: same length, same Bluestein block) go to one
// XCD, eight workgroups apart, so that they share that block in its L2.  A
// grid of 16k workgroups covers plans 0 .. 16k-1 once.
__device__ inline int pair_plan(int b)
{
  const int xcd = b & 7, k = b >> 3;
  return (((k >> 1) << 3) + xcd) * 2 + (k & 1);
}

// threads per ring: S/8 (8 values per thread and pass) within [256, 512];
// small rings still read all mmax+1 phases, which wants many threads
template <int S> constexpr int kThreads = (S / 8 < 256) ? 256 : ((S / 8 < 512) ? S / 8 : 512);

// Runs consume(i, load(i)) for i = lo+tid, lo+tid+NT, ... < hi with U loads
// in flight per thread (all U are issued before the first is used).
template <int NT, int U, typename L, typename C>
__device__ inline void batched(int tid, int lo, int hi, L &&load, C &&consume)
{
  for (int base = lo + tid; base < hi; base += NT * U) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * NT;
      v[u] = (i < hi) ? load(i) : make_double2(0., 0.);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * NT;
      if (i < hi) consume(i, v[u]);
    }
  }
}

} // namespace

// ===========================================================================
// synthesis: phase -> ring
// ===========================================================================
// minimum waves per SIMD: the radix-8 (equatorial) 4096 class, unstaged,
// runs two 512-thread workgroups per CU (72 KB of LDS each), 4 waves per
// SIMD, which caps its registers at 128 (it needs no more); so does the
// 4096 Bluestein class with the gathered two-alias fold (no staging area)
template <int S, bool BLUE, bool STAGED, int P2, int NA = 1>
constexpr int kMinBlocks = ((S == 4096 && !BLUE && !STAGED && P2 == 2) ||
                            (S == 4096 && BLUE && !STAGED && NA == 2)) ? 4 : 1;
// threads per workgroup: the unstaged 8192 Bluestein class (the polar
// rings) runs 1024 threads, 8 entries each: its 136 KB of LDS hold one
// workgroup per CU, and 16 waves hide more of the LDS latency of the three
// transforms than 8 do
template <int S, bool BLUE, bool STAGED>
constexpr int kSynThreads = (S == 8192 && BLUE && !STAGED) ? 1024 : kThreads<S>;

// NA (unstaged fold): most aliases m = k (mod n) of one bin in [0, mmax],
// 1 for rings with n > mmax, 2 for mmax < 2n
template <int S, bool BLUE, bool STAGED, int P2, typename T, int NA = 1>
__global__ void __launch_bounds__((kSynThreads<S, BLUE, STAGED>), (kMinBlocks<S, BLUE, STAGED, P2, NA>)) k_phase2ring_fft(RingArgs a, const RingFftPlan *plans,
                                                          int nplans, int comps_per_block, int vcap,
                                                          const double2 *btab)
{
  constexpr int NT = kSynThreads<S, BLUE, STAGED>, E = S / NT;
  // buf: the ring's FFT buffer; V (STAGED): up to vcap weighted, rotated phases
  extern __shared__ double2 smem[];
  const Lds buf{smem};
  double2 *V = smem + lds_pad(S);
  __shared__ double2 nyq;
  const int pi = pair_plan(blockIdx.x);
  if (pi >= nplans) return;
  const RingFftPlan &P = plans[pi];
  int tid = threadIdx.x;
  const int slot = P.ring >> 1, hemi = P.ring & 1;
  const RingDev r = a.ring[P.ring];
  const int n = P.n, N = P.N, mmax = a.mmax;
  const bool even = (n & 1) == 0;
  double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.weight : 1.;
  if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrtHalf;
  const bool norot = fabs(r.phi0) < 1e-14;
  const double2 *Bt = BLUE ? btab + P.boff : nullptr;
  // w e^{i m phi0} phase_m (rot = e^{i m phi0}), real part only for m = 0
  auto cval = [&](int m, double2 v, double2 rot) {
    v = cscale(v, wgt);
    if (m == 0) v.y = 0.;
    else if (!norot) v = cmul(v, rot);
    return v;
  };
  const double2 one = make_double2(1., 0.);
  // rotation steps: e^{i NT phi0} between a thread's bins, e^{i n phi0}
  const double2 rstep = uniform(norot ? one : expi_small((double)NT * r.phi0));
  const double2 rn = uniform(norot ? one : expi_small((double)n * r.phi0));
  // the same for every component (the fused fold and pack of even rings):
  // e^{i N phi0}; e^{i k phi0} of the thread's first bin k = tid and the
  // step 2 NT to its next pair of bins; the packing twiddle e^{2 pi i k/n}
  // of the first bin and the step NT between bins (where n is not a table
  // length).  Chained by products, they keep sine and cosine evaluations out
  // of the component loop.
  const double2 rN = uniform(norot ? one : expi_small((double)N * r.phi0));
  const double2 rk0 = norot ? one : expi_small((double)tid * r.phi0);
  const double2 rstep2 = uniform(norot ? one : expi_small((double)(2 * NT) * r.phi0));
  const double2 tw0 = expipi((double)tid * (2.0 / n));
  const double2 twstep = uniform(expipi((double)NT * (2.0 / n)));

  // bins: even n -> C_0..C_N, odd n -> C_0..C_{(n-1)/2}
  const int nbins = even ? N + 1 : (n + 1) / 2;
  int G = 1;
  while (G < 64 && 2 * G * nbins <= NT) G <<= 1;
  const int bins_per_round = NT / G;
  auto put_bin = [&](int k, double2 s, bool first) {
    if (even) {
      double2 &dst = (k < N) ? buf[k] : nyq;
      dst = first ? s : cadd(dst, s);
    } else {
      buf[k] = first ? s : cadd(buf[k], s);
      if (k > 0) buf[n - k] = conj(buf[k]);
    }
  };

  const int c0 = blockIdx.y * comps_per_block;
  const int c1 = min(c0 + comps_per_block, a.ncomp);
  for (int comp = c0; comp < c1; ++comp) {
    // at 1024 threads the register cap is 128: an opaque tid per component
    // keeps the compiler from holding every tid-derived LDS address live
    // across the loop (which spills)
    if constexpr (NT == 1024) asm volatile("" : "+v"(tid));
    // 1. fold: C_k = sum over m = k (mod n) of c_m + conj(sum over m = -k
    //    (mod n), m > 0, of c_m)
    bool packed = false;
    if constexpr (!STAGED) {
      // mmax < NA n: a bin's aliases are its first NA chain entries m, m + n
      if (even) {
        // even n: fold and pack in one pass.  The thread of k in [0, N/2]
        // forms the bins C_k and C_{N-k} (each from the alias chains of k and
        // n-k, and of N-k and N+k, conjugated for the second of each) and
        // writes the packed Z_k and Z_{N-k} (step 2 below), with no LDS round
        // trip between the two steps.  Rotations: e^{i(n-k)phi0} = rn
        // conj(rk), e^{i(N-k)phi0} = rN conj(rk), e^{i(N+k)phi0} = rN rk,
        // times rn per step along a chain.
        auto phase_at = [&](int m) {
          return a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)];
        };
        const double2 zero = make_double2(0., 0.);
        // bin pairs per thread and round: two with one alias, else one
        constexpr int U = (NA == 1) ? 2 : 1;
        const double2 rstepU = (U == 2) ? rstep2 : rstep;
        double2 rkb = rk0, twk = tw0;
        for (int base = tid; 2 * base <= N; base += U * NT) {
          double2 va[U][NA], vb[U][NA], vc[U][NA], vd[U][NA];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int k = base + u * NT, j = N - k, mb = (k == 0) ? n : n - k;
            const bool ok = 2 * k <= N;
#pragma unroll
            for (int q = 0; q < NA; ++q) {
              const int o = q * n;
              va[u][q] = (ok && k + o <= mmax) ? phase_at(k + o) : zero;
              vb[u][q] = (ok && mb + o <= mmax) ? phase_at(mb + o) : zero;
              vc[u][q] = (ok && j + o <= mmax) ? phase_at(j + o) : zero;
              vd[u][q] = (ok && N + k + o <= mmax) ? phase_at(N + k + o) : zero;
            }
          }
          double2 rk = rkb;
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int k = base + u * NT, j = N - k, mb = (k == 0) ? n : n - k;
            const double2 tw = (kFixedLen<S, P2> && !BLUE) ? pack_twiddle<S, P2, BLUE>(k, n) : twk;
            twk = cmul(twk, twstep);
            if (2 * k <= N) {
              double2 ra = rk, rb = (k == 0) ? rn : cmul(rn, conj(rk));
              double2 rc = cmul(rN, conj(rk)), rd = cmul(rN, rk);
              double2 ck = zero, cj = zero;
              if (k <= mmax) ck = cval(k, va[u][0], ra);
              if (mb <= mmax) ck = cadd(ck, conj(cval(mb, vb[u][0], rb)));
              if (j <= mmax) cj = cval(j, vc[u][0], rc);
              if (N + k <= mmax) cj = cadd(cj, conj(cval(N + k, vd[u][0], rd)));
#pragma unroll
              for (int q = 1; q < NA; ++q) {
                const int o = q * n;
                ra = cmul(ra, rn); rb = cmul(rb, rn); rc = cmul(rc, rn); rd = cmul(rd, rn);
                if (k + o <= mmax) ck = cadd(ck, cval(k + o, va[u][q], ra));
                if (mb + o <= mmax) ck = cadd(ck, conj(cval(mb + o, vb[u][q], rb)));
                if (j + o <= mmax) cj = cadd(cj, cval(j + o, vc[u][q], rc));
                if (N + k + o <= mmax) cj = cadd(cj, conj(cval(N + k + o, vd[u][q], rd)));
              }
              const double2 e = cadd(ck, conj(cj));
              const double2 o = cmul(csub(ck, conj(cj)), tw);
              buf[k] = make_double2(e.x - o.y, e.y + o.x);
              if (j < N && j != k) {
                const double2 e2 = cadd(cj, conj(ck));
                const double2 o2 = cmul(csub(cj, conj(ck)), make_double2(-tw.x, tw.y));
                buf[j] = make_double2(e2.x - o2.y, e2.y + o2.x);
              }
            }
            rk = cmul(rk, rstep);
          }
          rkb = cmul(rkb, rstepU);
        }
        packed = true;
      } else {
        // odd n: the chains of k and n-k (conjugated) per bin; 4 bins (8
        // loads) in flight per thread with one alias, 2 bins with two
        constexpr int U = (NA == 1) ? 4 : 2;
        for (int base = tid; base < nbins; base += U * NT) {
          double2 vp[U][NA], vn[U][NA];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int k = base + u * NT, mn = (k == 0) ? n : n - k;
#pragma unroll
            for (int q = 0; q < NA; ++q) {
              const int o = q * n;
              vp[u][q] = (k < nbins && k + o <= mmax)
                             ? a.phase[phase_index(comp, k + o, slot, hemi, a.nslots, a.nmb)]
                             : make_double2(0., 0.);
              vn[u][q] = (k < nbins && mn + o <= mmax)
                             ? a.phase[phase_index(comp, mn + o, slot, hemi, a.nslots, a.nmb)]
                             : make_double2(0., 0.);
            }
          }
          // e^{i k phi0} for this thread's bins; e^{i (n-k) phi0} from it
          double2 rk = norot ? one : expi_small((double)base * r.phi0);
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int k = base + u * NT, mn = (k == 0) ? n : n - k;
            if (k < nbins) {
              double2 rp = rk, rmn = (k == 0) ? rn : cmul(rn, conj(rk));
              double2 sum = make_double2(0., 0.);
              if (k <= mmax) sum = cval(k, vp[u][0], rp);
              if (mn <= mmax) sum = cadd(sum, conj(cval(mn, vn[u][0], rmn)));
#pragma unroll
              for (int q = 1; q < NA; ++q) {
                const int o = q * n;
                rp = cmul(rp, rn); rmn = cmul(rmn, rn);
                if (k + o <= mmax) sum = cadd(sum, cval(k + o, vp[u][q], rp));
                if (mn + o <= mmax) sum = cadd(sum, conj(cval(mn + o, vn[u][q], rmn)));
              }
              put_bin(k, sum, true);
            }
            rk = cmul(rk, rstep);
          }
        }
      }
      __syncthreads();
    } else {
      // stage vcap phases at a time in LDS, then sum each bin's aliases
      // with G lanes in a fixed order
      for (int m0 = 0; m0 <= mmax; m0 += vcap) {
        const int m1 = min(mmax + 1, m0 + vcap);
        for (int base = m0 + tid; base < m1; base += NT * 8) {
          double2 v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int m = base + u * NT;
            v[u] = (m < m1) ? a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)]
                            : make_double2(0., 0.);
          }
          double2 rot = norot ? one : expi_small((double)base * r.phi0);
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int m = base + u * NT;
            if (m < m1) V[m - m0] = cval(m, v[u], rot);
            rot = cmul(rot, rstep);
          }
        }
        __syncthreads();
        for (int kb = 0; kb < nbins; kb += bins_per_round) {
          const int k = kb + tid / G, part = tid % G;
          double2 sum = make_double2(0., 0.);
          if (k < nbins) {
            // first m >= m0 with m = k (mod n), then every G-th alias
            int64_t mp = k;
            if (mp < m0) mp += ((m0 - k + n - 1) / n) * (int64_t)n;
            for (int64_t m = mp + (int64_t)part * n; m < m1; m += (int64_t)G * n)
              sum = cadd(sum, V[m - m0]);
            // m = -k (mod n), m >= max(m0, 1)
            int64_t mn = (k == 0) ? n : n - k;
            const int lo = max(m0, 1);
            if (mn < lo) mn += ((lo - mn + n - 1) / n) * (int64_t)n;
            double2 t = make_double2(0., 0.);
            for (int64_t m = mn + (int64_t)part * n; m < m1; m += (int64_t)G * n)
              t = cadd(t, V[m - m0]);
            sum = cadd(sum, conj(t));
          }
          sum = group_sum(sum, G);
          if (part == 0 && k < nbins) put_bin(k, sum, m0 == 0);
        }
        __syncthreads();
      }
    }
  // 2. pack (even n): Z'_k = (C_k + C*_{N-k}) + i (C_k - C*_{N-k}) e^{2 pi i k/n}
    //    (e^{2 pi i (N-k)/n} = -conj(e^{2 pi i k/n}) since n = 2N)
    if (even && !packed) {
      for (int k = tid; 2 * k <= N; k += NT) {
        const int kp = N - k;
        const double2 ck = buf[k], ckp = (k == 0) ? nyq : buf[kp];
        const double2 tw = pack_twiddle<S, P2, BLUE>(k, n);
        const double2 e = cadd(ck, conj(ckp));
        const double2 o = cmul(csub(ck, conj(ckp)), tw);
        buf[k] = make_double2(e.x - o.y, e.y + o.x);
        if (kp < N && kp != k) {
          const double2 e2 = cadd(ckp, conj(ck));
          const double2 o2 = cmul(csub(ckp, conj(ck)), make_double2(-tw.x, tw.y));
          buf[kp] = make_double2(e2.x - o2.y, e2.y + o2.x);
        }
      }
      __syncthreads();
    }
    // 3. inverse FFT of length N
    ring_dft<NT, E, S, BLUE, P2>(buf, P, +1, Bt);
    // 4. store
    void *mp = a.maps.p[comp];
    const bool add = (a.flags & SHARP_ADD) != 0;
    const bool pairs = even && pair_aligned<T>(mp, r.ofs - a.map_base, r.stride);
    for (int j = tid; j < N; j += NT) {
      const double2 z = buf[j];
      if (even) {
        const int64_t i0 = r.ofs + (int64_t)(2 * j) * r.stride - a.map_base;
        const int64_t i1 = i0 + r.stride;
        if (pairs) {
          const double2 o = add ? ld_pair<T>(mp, i0) : make_double2(0., 0.);
          st_pair<T>(mp, i0, make_double2(o.x + z.x, o.y + z.y));
        } else {
          st<T>(mp, i0, (add ? ld<T>(mp, i0) : 0.) + z.x);
          st<T>(mp, i1, (add ? ld<T>(mp, i1) : 0.) + z.y);
        }
      } else {
        const int64_t i0 = r.ofs + (int64_t)j * r.stride - a.map_base;
        st<T>(mp, i0, (add ? ld<T>(mp, i0) : 0.) + z.x);
      }
    }
    __syncthreads();
  }
}

// ===========================================================================
// analysis: ring -> phase
// ===========================================================================
// threads per workgroup: 1024 for the 8192 Bluestein class, as in synthesis
// (115 VGPRs with the thread index opaque per component)
template <int S, bool BLUE>
constexpr int kAnaThreads = (S == 8192 && BLUE) ? 1024 : kThreads<S>;

template <int S, bool BLUE, int P2, typename T>
__global__ void __launch_bounds__((kAnaThreads<S, BLUE>), (kMinBlocks<S, BLUE, false, P2>)) k_ring2phase_fft(RingArgs a, const RingFftPlan *plans,
                                                          int nplans, int comps_per_block, int,
                                                          const double2 *btab)
{
  constexpr int NT = kAnaThreads<S, BLUE>, E = S / NT;
  extern __shared__ double2 smem[];
  const Lds buf{smem};
  __shared__ double2 nyq;
  const int pi = pair_plan(blockIdx.x);
  if (pi >= nplans) return;
  const RingFftPlan &P = plans[pi];
  int tid = threadIdx.x;
  const int slot = P.ring >> 1, hemi = P.ring & 1;
  const RingDev r = a.ring[P.ring];
  const int n = P.n, N = P.N, mmax = a.mmax;
  const bool even = (n & 1) == 0;
  double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.weight : 1.;
  if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrt2;
  const bool norot = fabs(r.phi0) < 1e-14;

  const double2 *Bt = BLUE ? btab + P.boff : nullptr;
  const double2 one = make_double2(1., 0.);
  // the same for every component: e^{-i m phi0} of the thread's first order
  // m = tid and the steps NT (between its orders) and 8 NT (a product with
  // the step of 8 orders every 8 orders keeps the chain short); the unpacking
  // twiddle e^{2 pi i k/n} of the first bin k = tid and the step NT (where n
  // is not a table length).  Chained by products, they keep sine and cosine
  // evaluations out of the component loop.
  const double2 rot0 = norot ? one : expi_small((double)tid * (-r.phi0));
  const double2 rstep = uniform(norot ? one : expi_small((double)NT * (-r.phi0)));
  const double2 rstep8 = uniform(norot ? one : expi_small((double)(8 * NT) * (-r.phi0)));
  const double2 tw0 = expipi((double)tid * (2.0 / n));
  const double2 twstep = uniform(expipi((double)NT * (2.0 / n)));

  const int c0 = blockIdx.y * comps_per_block;
  const int c1 = min(c0 + comps_per_block, a.ncomp);
  for (int comp = c0; comp < c1; ++comp) {
    // keeps tid-derived addresses out of the loop-carried registers
    if constexpr (NT == 1024) asm volatile("" : "+v"(tid));
    const void *mp = a.maps.p[comp];
    // pixel pairs as one vector access: unit stride, pair start aligned
    const bool pairs = even && pair_aligned<T>(mp, r.ofs - a.map_base, r.stride);
    // 1. load: even n -> z_j = f_2j + i f_2j+1, odd n -> z_j = f_j
    batched<NT, (P2 == 2) ? 4 : 8>(tid, 0, N,
        [&](int j) {
          if (even) {
            const int64_t i0 = r.ofs + (int64_t)(2 * j) * r.stride - a.map_base;
            if (pairs) return ld_pair<T>(mp, i0);
            return make_double2(ld<T>(mp, i0), ld<T>(mp, i0 + r.stride));
          }
          const int64_t i0 = r.ofs + (int64_t)j * r.stride - a.map_base;
          return make_double2(ld<T>(mp, i0), 0.);
        },
        [&](int j, double2 v) { buf[j] = v; });
    __syncthreads();
    // 2. forward FFT
    ring_dft<NT, E, S, BLUE, P2>(buf, P, -1, Bt);
    // 3. unpack (even n): X_k = (Z_k + Z*_{N-k})/2 - i/2 e^{-2 pi i k/n} (Z_k - Z*_{N-k})
    //    (e^{-2 pi i (N-k)/n} = -e^{2 pi i k/n} since n = 2N)
    if (even) {
      double2 twk = tw0;
      for (int k = tid; 2 * k <= N; k += NT) {
        const int kp = N - k;
        const double2 tw = (kFixedLen<S, P2> && !BLUE) ? pack_twiddle<S, P2, BLUE>(k, n) : twk;
        twk = cmul(twk, twstep);
        if (k == 0) {
          const double2 z0 = buf[0];
          buf[0] = make_double2(z0.x + z0.y, 0.);
          nyq = make_double2(z0.x - z0.y, 0.);
        } else {
          const double2 zk = buf[k], zkp = buf[kp];
          const double2 s1 = cscale(cadd(zk, conj(zkp)), 0.5);
          const double2 d1 = cmul(cscale(csub(zk, conj(zkp)), 0.5), conj(tw));
          buf[k] = make_double2(s1.x + d1.y, s1.y - d1.x);
          if (kp != k) {
            const double2 s2 = cscale(cadd(zkp, conj(zk)), 0.5);
            const double2 d2 = cmul(cscale(csub(zkp, conj(zk)), 0.5),
                                    make_double2(-tw.x, -tw.y));
            buf[kp] = make_double2(s2.x + d2.y, s2.y - d2.x);
          }
        }
      }
      __syncthreads();
    }
    // 4. phases
    double2 rot = rot0, rot8 = rot0;
    int step = 0;
    for (int m = tid; m <= mmax; m += NT, ++step) {
      const int k = m % n;
      double2 v;
      if (even) v = (k < N) ? buf[k] : ((k == N) ? nyq : conj(buf[n - k]));
      else v = buf[k];
      v = cscale(v, wgt);
      if (!norot) {
        // e^{-i m phi0}: products by the step of 8 orders every 8 orders,
        // by the step of one order between
        if (step > 0) {
          if ((step & 7) == 0) rot = rot8 = cmul(rot8, rstep8);
          else rot = cmul(rot, rstep);
        }
        v = cmul(v, rot);
      }
      a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)] = v;
    }
    __syncthreads();
  }
}

// ===========================================================================
// launchers: one launch per LDS size class
// ===========================================================================
#define HIPCHK(x)                                                          \
  do {                                                                     \
    const hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) SHP_FAIL(hipGetErrorString(e_));                 \
  } while (0)

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

// Bluestein kernel spectra for the analysis sign, one plan per workgroup
template <int S>
__global__ void __launch_bounds__(kThreads<S>) k_bluestein_table(const RingFftPlan *plans,
                                                           double2 *btab)
{
  constexpr int NT = kThreads<S>, E = S / NT;
  extern __shared__ double2 smem[];
  const RingFftPlan &P = plans[blockIdx.x];
  const Lds buf{smem};
  bluestein_kernel<NT, E>(buf, P, -1);
  // halves of the symmetric sequences (ring_dft): B_0..B_{M/2}, c_0..c_{N/2}
  for (int j = threadIdx.x; 2 * j <= P.M; j += NT) btab[P.boff + j] = buf[j];
  for (int j = threadIdx.x; 2 * j <= P.N; j += NT)
    btab[P.boff + P.M / 2 + 1 + j] = chirp(j, P.N, -1);
}

// raises a kernel's dynamic LDS limit once per new maximum (callers hold the
// library's lock)
template <typename K> static void set_lds(K kern, size_t lds)
{
  static std::map<const void *, size_t> done;
  const void *k = reinterpret_cast<const void *>(kern);
  if (lds <= 65536 || done[k] >= lds) return;
  HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  done[k] = lds;
}

static size_t plan_need(const RingFftPlan &p) { return p.bluestein ? p.M : p.N; }

// A plan run by the fixed passes of its class must have the class's length
// and the passes' radices (fft_plan builds exactly these; this guards the
// two against drifting apart).
template <int S, bool BLUE> static bool fixed_fft_plan_ok(const RingFftPlan &p)
{
  typedef FixedPasses<S, BLUE> FP;
  if ((BLUE ? p.M : p.N) != S || p.nrad != FP::n) return false;
  for (int q = 0; q < FP::n; ++q)
    if (p.rad[q] != FP::r[q]) return false;
  return true;
}

template <int S, int P2>
static void launch_range(bool synth, bool blue, bool staged, const RingArgs &a,
                         const RingFftPlan *plans, const RingFftPlan *hplans, int nplans,
                         const double2 *btab, hipStream_t st, int na = 1)
{
  if (nplans <= 0) return;
  // split components over blocks only when there are few rings
  int cpb = a.ncomp;
  while (cpb > 1 && (int64_t)nplans * ((a.ncomp + cpb - 1) / cpb) < 2048) cpb = (cpb + 1) / 2;
  const dim3 grid((nplans + 15) / 16 * 16, (a.ncomp + cpb - 1) / cpb);
  size_t need = 0;
  for (int i = 0; i < nplans; ++i) need = std::max(need, plan_need(hplans[i]));
  if constexpr (kFixedLen<S, P2>) {
    for (int i = 0; i < nplans; ++i)
      if (!(blue ? fixed_fft_plan_ok<S, true>(hplans[i]) : fixed_fft_plan_ok<S, false>(hplans[i])))
        SHP_FAIL("ring FFT: a plan does not match the fixed passes of its class");
  }
  // the staging area V starts after the class's S-entry buffer (padded)
  const size_t fbytes = (size_t)lds_pad(staged ? S : (int)need) * sizeof(double2);
  // staged synthesis keeps up to vcap phases next to the FFT buffer: the
  // whole ring's mmax+1 when the LDS allows, else it loops
  int vcap = 0;
  if (staged) {
    const size_t budget = (S >= 4096) ? (size_t)a.max_lds - 256
                                      : std::min<size_t>(81920, (size_t)a.max_lds - 256);
    vcap = (int)std::min<size_t>((size_t)a.mmax + 1, (budget - fbytes) / sizeof(double2));
    if (vcap < 1) SHP_FAIL("ring FFT: no LDS left for the phase staging");
  }
  const size_t lds = fbytes + (size_t)vcap * sizeof(double2);
  auto go = [&](auto kern, int nt = kThreads<S>) {
    set_lds(kern, lds);
    hipLaunchKernelGGL(kern, grid, dim3(nt), lds, st, a, plans, nplans, cpb, vcap, btab);
  };
#define SHP_SYN(B, ST)                                                     \
  do {                                                                     \
    constexpr int nt = kSynThreads<S, B, ST>;                              \
    if constexpr (!ST && S >= 2048) {                                      \
      if (na == 2) {                                                       \
        if (a.dp) go(k_phase2ring_fft<S, B, false, P2, double, 2>, nt);    \
        else go(k_phase2ring_fft<S, B, false, P2, float, 2>, nt);          \
        break;                                                             \
      }                                                                    \
    }                                                                      \
    if (a.dp) go(k_phase2ring_fft<S, B, ST, P2, double>, nt);              \
    else go(k_phase2ring_fft<S, B, ST, P2, float>, nt);                    \
  } while (0)
  // Bluestein lengths are powers of two: no Bluestein kernel of the radix
  // set with 3, 5, 7
  if constexpr (P2 == 0) {
    if (blue) SHP_FAIL("ring FFT: Bluestein plan with an odd radix");
  }
  if (synth) {
    if constexpr (P2 > 0) {
      if (blue) {
        if (staged) SHP_SYN(true, true);
        else SHP_SYN(true, false);
      }
    }
    if (!blue) {
      if (staged) SHP_SYN(false, true);
      else SHP_SYN(false, false);
    }
  } else {
    if constexpr (P2 > 0) {
      if (blue) {
        if (a.dp) go(k_ring2phase_fft<S, true, P2, double>, kAnaThreads<S, true>);
        else go(k_ring2phase_fft<S, true, P2, float>, kAnaThreads<S, true>);
      }
    }
    if (!blue) {
      if (a.dp) go(k_ring2phase_fft<S, false, P2, double>);
      else go(k_ring2phase_fft<S, false, P2, float>);
    }
  }
#undef SHP_SYN
  check_launch(synth ? "phase2ring_fft" : "ring2phase_fft");
}

template <int S, int P2>
static void launch_class(bool synth, bool blue, const RingArgs &a, const RingFftPlan *plans,
                         const RingFftPlan *hplans, int nplans, const double2 *btab,
                         hipStream_t st)
{
  if (!synth) {
    launch_range<S, P2>(false, blue, false, a, plans, hplans, nplans, btab, st);
    return;
  }
  // plans are sorted by decreasing n: rings with n > mmax fold directly,
  // those with 2n > mmax (in the classes of 2048 entries and more) gather
  // their two aliases per bin from the phase array, the rest stage their
  // phases in LDS
  int k = 0;
  while (k < nplans && hplans[k].n > a.mmax) ++k;
  int k2 = k;
  if (S >= 2048)
    while (k2 < nplans && 2 * (int64_t)hplans[k2].n > a.mmax) ++k2;
  launch_range<S, P2>(true, blue, false, a, plans, hplans, k, btab, st);
  launch_range<S, P2>(true, blue, false, a, plans + k, hplans + k, k2 - k, btab, st, 2);
  launch_range<S, P2>(true, blue, true, a, plans + k2, hplans + k2, nplans - k2, btab, st);
}

template <int P2>
static void launch_kind(int c, bool synth, bool blue, const RingArgs &a,
                        const RingFftPlan *p, const RingFftPlan *hp, int np,
                        const double2 *btab, hipStream_t st)
{
  switch (c) {
  case 0: launch_class<512, P2>(synth, blue, a, p, hp, np, btab, st); break;
  case 1: launch_class<1024, P2>(synth, blue, a, p, hp, np, btab, st); break;
  case 2: launch_class<2048, P2>(synth, blue, a, p, hp, np, btab, st); break;
  case 3: launch_class<4096, P2>(synth, blue, a, p, hp, np, btab, st); break;
  case 4: launch_class<8192, P2>(synth, blue, a, p, hp, np, btab, st); break;
  default: break;
  }
}

// ===========================================================================
// Rings too long for every LDS class (n > 8192 complex values, or a Bluestein
// length above 8192): the same sums as the direct-sum kernels,
//   synthesis  f_j = Re sum_k S_k e^{+2 pi i jk/n},  S_k = sum_{m = k mod n} c_m,
//   analysis   phase_m = w e^{-i m phi0} F_{m mod n},  F_k = sum_j f_j e^{-2 pi i jk/n},
// with one complex FFT of the ring's full length n: radix-16/8/4/2 Stockham
// passes through global memory when n is a power of two, otherwise
// Bluestein's chirp-z form with a power-of-two convolution length L >= 2n-1
// (jk = (j^2 + k^2 - (k-j)^2) / 2):
//   F_k = c_k sum_j (f_j c_j) b_{k-j},  c_j = e^{-i pi j^2/n},  b_d = e^{+i pi d^2/n},
// the convolution as FFT^+(FFT^-(a) . B) with B = FFT^-(b) / L precomputed per
// ring length (b is even, so the synthesis, whose chirps are the conjugates,
// uses conj(B)).  One workgroup takes one (ring, component) at a time and
// keeps its data in its own slot of global scratch (two buffers of the
// largest L); passes are separated by a device-scope fence and a barrier.
// ===========================================================================
namespace {

constexpr int kLongThreads = 512;

__device__ inline void long_sync()
{
  __threadfence();
  __syncthreads();
}

// c_j = e^{-i pi j^2 / n}, the angle reduced exactly (j^2 mod 2n)
__device__ inline double2 long_chirp(int64_t j, int n)
{
  const int64_t q = (j * j) % (2 * (int64_t)n);
  return expipi(-(double)q / (double)n);
}

// One Stockham pass of radix R over a length-L sequence, src -> dst; p is the
// product of the radices of the earlier passes.  Butterfly i takes
// src[i + r L/R], multiplies entry r by e^{sign 2 pi i r k/(pR)} (k = i mod
// p), transforms, and writes output q to dst[(i-k) R + k + q p].
template <int R>
__device__ inline void long_pass(const double2 *src, double2 *dst, int L, int p, int sign)
{
  const int T = L / R;
  for (int i = threadIdx.x; i < T; i += kLongThreads) {
    const int k = i & (p - 1);
    double2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = src[i + r * T];
    if (p > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r)
        u[r] = cmul(u[r], expipi(sign * 2.0 * (double)(r * k) / (double)(p * R)));
    }
    Dft<R>::run(u, sign);
    const int j = (i - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[j + r * p] = u[r];
  }
}

// FFT of length L = 2^lg of x (sum_j x_j e^{sign 2 pi i jk/L}); y is scratch
// of L entries; returns the buffer that holds the result
__device__ inline double2 *long_fft(double2 *x, double2 *y, int lg, int sign)
{
  const int L = 1 << lg;
  double2 *src = x, *dst = y;
  int p = 1, rem = lg;
  while (rem > 0) {
    const int b = rem >= 4 ? 4 : rem;
    if (b == 4) long_pass<16>(src, dst, L, p, sign);
    else if (b == 3) long_pass<8>(src, dst, L, p, sign);
    else if (b == 2) long_pass<4>(src, dst, L, p, sign);
    else long_pass<2>(src, dst, L, p, sign);
    long_sync();
    double2 *t = src;
    src = dst;
    dst = t;
    p <<= b;
    rem -= b;
  }
  return src;
}

// the Bluestein convolution of the chirped sequence in x (L entries):
// FFT^+(FFT^-(x) . B), B or conj(B); returns the buffer with the result
__device__ inline double2 *long_convolve(double2 *x, double2 *y, int lg, const double2 *B,
                                         bool conjB)
{
  const int L = 1 << lg;
  double2 *f = long_fft(x, y, lg, -1);
  for (int k = threadIdx.x; k < L; k += kLongThreads) {
    const double2 b = B[k];
    f[k] = cmul(f[k], conjB ? conj(b) : b);
  }
  long_sync();
  return long_fft(f, f == x ? y : x, lg, +1);
}

template <typename T>
__global__ void __launch_bounds__(kLongThreads) k_long_synth(RingArgs a, LongRingArgs la)
{
  double2 *X = la.scratch + (int64_t)blockIdx.x * la.slot;
  const int nitems = la.nplans * a.ncomp;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const LongRingPlan pl = la.plans[it / a.ncomp];
    const int comp = it % a.ncomp;
    const int slot = pl.ring >> 1, hemi = pl.ring & 1;
    const RingDev r = a.ring[pl.ring];
    const int n = pl.n, L = 1 << pl.lg;
    double2 *Y = X + L;
    double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.weight : 1.;
    if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrtHalf;
    const bool norot = fabs(r.phi0) < 1e-14;
    // S_k: the weighted, rotated phases folded onto bin k = m mod n (factor 2
    // for m > 0, the real part only for m = 0), then the conjugate chirp
    for (int k = threadIdx.x; k < L; k += kLongThreads) {
      double2 v = make_double2(0., 0.);
      if (k < n) {
        for (int m = k; m <= a.mmax; m += n) {
          double2 c = cscale(a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)], wgt);
          if (m == 0) {
            c.y = 0.;
          } else {
            if (!norot) c = cmul(c, expi_small(m * r.phi0));
            c = cscale(c, 2.);
          }
          v = cadd(v, c);
        }
        if (pl.bluestein) v = cmul(v, conj(long_chirp(k, n)));
      }
      X[k] = v;
    }
    long_sync();
    const double2 *res = pl.bluestein ? long_convolve(X, Y, pl.lg, la.btab + pl.boff, true)
                                      : long_fft(X, Y, pl.lg, +1);
    void *mp = a.maps.p[comp];
    for (int j = threadIdx.x; j < n; j += kLongThreads) {
      double2 v = res[j];
      if (pl.bluestein) v = cmul(v, conj(long_chirp(j, n)));
      const int64_t idx = r.ofs + (int64_t)j * r.stride - a.map_base;
      const double old = (a.flags & SHARP_ADD) ? ld<T>(mp, idx) : 0.;
      st<T>(mp, idx, old + v.x);
    }
    long_sync();   // the slot is reused by the next item
  }
}

template <typename T>
__global__ void __launch_bounds__(kLongThreads) k_long_anal(RingArgs a, LongRingArgs la)
{
  double2 *X = la.scratch + (int64_t)blockIdx.x * la.slot;
  const int nitems = la.nplans * a.ncomp;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const LongRingPlan pl = la.plans[it / a.ncomp];
    const int comp = it % a.ncomp;
    const int slot = pl.ring >> 1, hemi = pl.ring & 1;
    const RingDev r = a.ring[pl.ring];
    const int n = pl.n, L = 1 << pl.lg;
    double2 *Y = X + L;
    double wgt = (a.flags & SHARP_USE_WEIGHTS) ? r.weight : 1.;
    if (a.flags & SHARP_REAL_HARMONICS) wgt *= kSqrt2;
    const void *mp = a.maps.p[comp];
    for (int j = threadIdx.x; j < L; j += kLongThreads) {
      double2 v = make_double2(0., 0.);
      if (j < n) {
        v.x = ld<T>(mp, r.ofs + (int64_t)j * r.stride - a.map_base);
        if (pl.bluestein) v = cscale(long_chirp(j, n), v.x);
      }
      X[j] = v;
    }
    long_sync();
    const double2 *res = pl.bluestein ? long_convolve(X, Y, pl.lg, la.btab + pl.boff, false)
                                      : long_fft(X, Y, pl.lg, -1);
    const bool norot = fabs(r.phi0) < 1e-14;
    for (int m = threadIdx.x; m <= a.mmax; m += kLongThreads) {
      const int k = m % n;
      double2 f = res[k];
      if (pl.bluestein) f = cmul(f, long_chirp(k, n));
      f = cscale(f, wgt);
      if (!norot) f = cmul(f, expi_small(-m * r.phi0));
      a.phase[phase_index(comp, m, slot, hemi, a.nslots, a.nmb)] = f;
    }
    long_sync();
  }
}

// B = FFT^-(b) / L for each plan (one per Bluestein length n):
// b_d = e^{+i pi d^2/n} at d and L-d for 0 <= d < n, zero elsewhere
__global__ void __launch_bounds__(kLongThreads) k_long_table(const LongRingPlan *plans, int nplans,
                                                             double2 *scratch, int64_t slot,
                                                             double2 *btab)
{
  double2 *X = scratch + (int64_t)blockIdx.x * slot;
  for (int it = blockIdx.x; it < nplans; it += gridDim.x) {
    const LongRingPlan pl = plans[it];
    const int n = pl.n, L = 1 << pl.lg;
    for (int d = threadIdx.x; d < L; d += kLongThreads) {
      double2 v = make_double2(0., 0.);
      if (d < n) v = conj(long_chirp(d, n));
      else if (d > L - n) v = conj(long_chirp(L - d, n));
      X[d] = v;
    }
    long_sync();
    const double2 *f = long_fft(X, X + L, pl.lg, -1);
    const double s = 1.0 / (double)L;
    for (int k = threadIdx.x; k < L; k += kLongThreads) btab[pl.boff + k] = cscale(f[k], s);
    long_sync();
  }
}

} // namespace

void launch_long_rings(bool synth, const RingArgs &a, const LongRingArgs &la, int nslots_wg,
                       hipStream_t st)
{
  if (la.nplans <= 0) return;
  const dim3 grid(nslots_wg), block(kLongThreads);
  if (synth) {
    if (a.dp) hipLaunchKernelGGL(k_long_synth<double>, grid, block, 0, st, a, la);
    else hipLaunchKernelGGL(k_long_synth<float>, grid, block, 0, st, a, la);
  } else {
    if (a.dp) hipLaunchKernelGGL(k_long_anal<double>, grid, block, 0, st, a, la);
    else hipLaunchKernelGGL(k_long_anal<float>, grid, block, 0, st, a, la);
  }
  check_launch(synth ? "long-ring synthesis" : "long-ring analysis");
}

void launch_long_table(const LongRingPlan *plans, int nplans, double2 *scratch, int64_t slot,
                       int nslots_wg, double2 *btab, hipStream_t st)
{
  if (nplans <= 0) return;
  hipLaunchKernelGGL(k_long_table, dim3(nslots_wg), dim3(kLongThreads), 0, st, plans, nplans,
                     scratch, slot, btab);
  check_launch("long-ring Bluestein table");
}

void launch_ring_fft(bool synth, const RingArgs &a, const FftClasses &fc, hipStream_t st0,
                     const FftFork *fk)
{
  // (fk) the classes go round-robin over st0 and fk's two streams, which
  // start after everything before on st0 and which st0 waits for at the end
  if (fk) {
    HIPCHK(hipEventRecord(fk->fork, st0));
    for (int i = 0; i < 2; ++i) HIPCHK(hipStreamWaitEvent(fk->aux[i], fk->fork, 0));
  }
  int next = 0;
  // largest rings first: they take longest
  for (int c = kNumFftClasses - 1; c >= 0; --c) {
    for (int kind = kFftKinds - 1; kind >= 0; --kind) {
      const int idx = kFftKinds * c + kind;
      if (fc.count[idx] == 0) continue;
      const hipStream_t st = (!fk || next == 0) ? st0 : fk->aux[next - 1];
      next = (next + 1) % 3;
      const RingFftPlan *p = fc.plans[idx], *hp = fc.hplans[idx];
      const int np = fc.count[idx];
      const bool blue = (kind & 1) != 0;
      switch (kind >> 1) {
      case 2: launch_kind<2>(c, synth, blue, a, p, hp, np, fc.btab, st); break;
      case 1: launch_kind<1>(c, synth, blue, a, p, hp, np, fc.btab, st); break;
      default: launch_kind<0>(c, synth, blue, a, p, hp, np, fc.btab, st); break;
      }
    }
  }
  if (fk)
    for (int i = 0; i < 2; ++i) {
      HIPCHK(hipEventRecord(fk->join[i], fk->aux[i]));
      HIPCHK(hipStreamWaitEvent(st0, fk->join[i], 0));
    }
}

template <int S>
static void table_class(const RingFftPlan *plans, const RingFftPlan *hplans, int np,
                        double2 *btab, hipStream_t st)
{
  if (np <= 0) return;
  size_t need = 0;
  for (int i = 0; i < np; ++i) need = std::max(need, plan_need(hplans[i]));
  const size_t lds = (size_t)lds_pad((int)need) * sizeof(double2);
  set_lds(k_bluestein_table<S>, lds);
  hipLaunchKernelGGL(k_bluestein_table<S>, dim3(np), dim3(kThreads<S>), lds, st, plans, btab);
  check_launch("bluestein_table");
}

// the roots of unity of g_twiddle, exact angles 2 j / kTwN (in units of pi)
__global__ void __launch_bounds__(256) k_twiddle_init()
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < kTwN) {
    double sn, cs;
    sincospi(2.0 * j / kTwN, &sn, &cs);
    g_twiddle[j] = make_double2(cs, sn);
  }
}

void ensure_twiddles(hipStream_t st)
{
  // once per device, before any ring FFT on it (callers hold the library's
  // lock); finished before the caller goes on, so that transforms on other
  // streams find the table written
  static bool done[256] = {};
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 256) SHP_FAIL("ring FFT: device index out of range");
  if (done[dev]) return;
  hipLaunchKernelGGL(k_twiddle_init, dim3(kTwN / 256), dim3(256), 0, st);
  check_launch("twiddle table");
  HIPCHK(hipStreamSynchronize(st));
  done[dev] = true;
}

void launch_bluestein_table(const FftClasses &u, double2 *btab, hipStream_t st)
{
  for (int c = 0; c < kNumFftClasses; ++c) {
    const int idx = kFftKinds * c + 1;   // the table list holds every Bluestein plan here
    const RingFftPlan *p = u.plans[idx], *hp = u.hplans[idx];
    const int np = u.count[idx];
    switch (c) {
    case 0: table_class<512>(p, hp, np, btab, st); break;
    case 1: table_class<1024>(p, hp, np, btab, st); break;
    case 2: table_class<2048>(p, hp, np, btab, st); break;
    case 3: table_class<4096>(p, hp, np, btab, st); break;
    case 4: table_class<8192>(p, hp, np, btab, st); break;
    default: break;
    }
  }
}

} // namespace shp

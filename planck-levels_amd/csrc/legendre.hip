// legendre.hip -- FP64 Legendre stage of the SHT on CDNA4 (gfx950).
//
// Synthesis (a_lm -> phase): one wavefront per (m, 64 ring pairs).  Each lane
// owns one ring pair and runs the three-term recurrence over l in registers;
// the per-m recurrence coefficients and the a_lm of that m are wave-uniform
// and are fetched through the scalar data path (s_load), so the vector ALUs
// only see the recurrence and the accumulation.  A workgroup is MB=8
// wavefronts with consecutive m over the same 64 ring pairs, which fills the
// 128-byte phase blocks phase[comp][m/8][slot][hemi][0..7].
//
// Analysis (phase -> a_lm): one workgroup (4 wavefronts) per m, looping over
// the ring pairs 256 at a time.  Each lane accumulates lambda_lm(ring) *
// phase(ring) for 8 consecutive l, the wavefront reduces the 8 l-values
// across its 64 lanes with a transpose-reduce (each exchange halves the
// number of values a lane carries), the 4 wavefronts combine in LDS and one
// fixed thread per output adds into the canonical a_lm row of this m, so the
// summation order is deterministic.
//
// Exponent scaling follows libsharp: start values sin^m(theta) are formed
// with a separate power-of-2 exponent, the recurrence runs in the scaled
// domain until some lane reaches the IEEE range, then with per-lane
// correction factors until every lane has, then unscaled.  "Some lane" /
// "every lane" are wavefront ballots here where libsharp uses its SIMD
// block (libsharp/sharp_core_inc.c:146-207, sharp_core_inc2.c:164-280).
//
// Spin s > 0 uses the pair of Wigner-d recurrences lambda_plus / lambda_minus
// (libsharp/sharp_core_inc.c:209-293, sharp_ylmgen_c.c:162-197) and the E/B ->
// Q/U combination of libsharp/sharp_core_inc2.c:282-430; ALM2MAP_DERIV1 is the
// same machinery with the curl a_lm identically zero
// (libsharp/sharp_core_inc2.c:521-619).

#include <hip/hip_runtime.h>
#include <type_traits>
#include "kernels.hpp"
#include "legendre_common.hpp"

namespace shp {

// ---------------------------------------------------------------------------
// transpose-reduce over the 64 lanes of a wavefront.  On entry every lane
// holds C values; on exit v[0] of lane L holds the sum over all lanes of value
// number xred_index<C>(L), replicated over the 64/C lanes that differ only in
// their low log2(64/C) bits.
// ---------------------------------------------------------------------------
template <int C, int D> struct XRed {
  static __device__ inline void run(double *v, int lane)
  {
    if constexpr (D == 0) {
      return;
    } else if constexpr (C > 1) {
      const bool up = (lane & D) != 0;
#pragma unroll
      for (int i = 0; i < C / 2; ++i) {
        const double send = up ? v[i] : v[i + C / 2];
        const double keep = up ? v[i + C / 2] : v[i];
        v[i] = keep + __shfl_xor(send, D);
      }
      XRed<C / 2, D / 2>::run(v, lane);
    } else {
      v[0] += __shfl_xor(v[0], D);
      XRed<1, D / 2>::run(v, lane);
    }
  }
};

template <int C> __device__ inline int xred_index(int lane)
{
  int idx = 0, c = C;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    if (c > 1) {
      if (lane & d) idx += c / 2;
      c /= 2;
    }
  }
  return idx;
}

// ===========================================================================
// spin 0 synthesis
// ===========================================================================
// One wavefront: order m (wave-uniform) for the 64 ring pairs of slot block
// sb, TG transforms from a.t0.
// CROW: rows through the constant address space (the batch kernels, whose
// job arguments come from memory; kernel-argument pointers get scalar loads
// without it, and plain pointers schedule slightly better)
template <int TG, bool CROW = false>
__device__ void synth_spin0_tile(const LegArgs &a, int m, int sb)
{
  const int lane = threadIdx.x;
  if (m > a.mmax) return;
  const int lmax = a.lmax;
  const int slot = sb * 64 + lane;
  const bool valid = slot < a.nslots;
  const int s = valid ? slot : a.nslots - 1;
  const double cth = a.cth[s], sth = a.sth[s];
  const bool active = valid && (a.mlim[s] >= m);

  double p1r[TG], p1i[TG], p2r[TG], p2i[TG];
#pragma unroll
  for (int t = 0; t < TG; ++t) p1r[t] = p1i[t] = p2r[t] = p2i[t] = 0.;

  if (wave_any(active)) {
    typedef std::conditional_t<CROW, URow2, const double2 *> Row2;
    const Row2 rf = Row2(reinterpret_cast<const double2 *>(a.tab) + (tab_ofs(m, lmax) - m));
    const int nc = a.ncomp_alm;
    const Row2 al = Row2(a.alm + ((tri_ofs(m, lmax) - m) * nc + a.t0));

    double lam1 = 0., lam2;
    int scale;
    mypow(sth, m, lam2, scale);
    const double mf = a.mfac[m];
    lam2 *= (m & 1) ? -mf : mf;
    normalize(lam2, scale, kFTol);

    int l = m;
    bool dead = false;
    // below the IEEE range on every lane: advance without accumulating
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
      bool done = false;
      // some lanes still scaled: accumulate with correction factors
      while (wave_any(scale < 0)) {
        {
          const double tmp = lam2 * cf;
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            const double2 x = al[(int64_t)l * nc + t];
            p1r[t] += tmp * x.x; p1i[t] += tmp * x.y;
          }
        }
        if (++l > lmax) { done = true; break; }
        {
          const double2 r = rf[l - 1];
          lam1 = (cth * lam2) * r.x - lam1 * r.y;
          const double tmp = lam1 * cf;
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            const double2 x = al[(int64_t)l * nc + t];
            p2r[t] += tmp * x.x; p2i[t] += tmp * x.y;
          }
        }
        if (++l > lmax) { done = true; break; }
        const double2 r = rf[l - 1];
        lam2 = (cth * lam1) * r.x - lam2 * r.y;
        if (fabs(lam2) > kFTol) {
          lam1 *= kFSmall; lam2 *= kFSmall; ++scale;
          cf = corfac(scale);
        }
      }
      if (!done) {
        lam1 *= cf;
        lam2 *= cf;
        // hot loop: 3 FP64 ops of recurrence + 4*TG of accumulation per l
        // pair.  The next pair's coefficients and a_lm (scalar loads) are
        // requested right after the first use of the current ones (which
        // waits for every scalar load in flight), so they have a whole pair
        // to arrive; rows up to lmax+2 are readable (table padding, next m
        // or the slack rows of the canonical array).
        double2 r0 = rf[l], r1 = rf[l + 1];
        double2 x0[TG], x1[TG];
#pragma unroll
        for (int t = 0; t < TG; ++t) {
          x0[t] = al[(int64_t)l * nc + t];
          x1[t] = al[(int64_t)(l + 1) * nc + t];
        }
        while (l < lmax) {
          lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
          __builtin_amdgcn_sched_barrier(0);
          const double2 nr0 = rf[l + 2], nr1 = rf[l + 3];
          double2 nx0[TG], nx1[TG];
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            nx0[t] = al[(int64_t)(l + 2) * nc + t];
            nx1[t] = al[(int64_t)(l + 3) * nc + t];
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            p1r[t] = fma(lam2, x0[t].x, p1r[t]); p1i[t] = fma(lam2, x0[t].y, p1i[t]);
            p2r[t] = fma(lam1, x1[t].x, p2r[t]); p2i[t] = fma(lam1, x1[t].y, p2i[t]);
          }
          lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
          r0 = nr0;
          r1 = nr1;
#pragma unroll
          for (int t = 0; t < TG; ++t) { x0[t] = nx0[t]; x1[t] = nx1[t]; }
          l += 2;
        }
        if (l == lmax) {
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            p1r[t] += lam2 * x0[t].x; p1i[t] += lam2 * x0[t].y;
          }
        }
      }
    }
  }
  if (!valid) return;
  const bool pr = a.ispair[slot] != 0;
#pragma unroll
  for (int t = 0; t < TG; ++t) {
    double2 n = make_double2(0., 0.), sv = make_double2(0., 0.);
    if (active) {
      n = make_double2(p1r[t] + p2r[t], p1i[t] + p2i[t]);
      sv = make_double2(p1r[t] - p2r[t], p1i[t] - p2i[t]);
    }
    a.phase[phase_index(a.t0 + t, m, slot, 0, a.nslots, a.nmb)] = n;
    if (pr) a.phase[phase_index(a.t0 + t, m, slot, 1, a.nslots, a.nmb)] = sv;
  }
}

// ===========================================================================
// spin s synthesis (and ALM2MAP_DERIV1)
// ===========================================================================
struct QU { double qr, qi, ur, ui; };

// px += lw*(E,B), py += lx*rot(E,B)  (libsharp/sharp_core_inc2.c:282-306)
template <bool DERIV>
__device__ inline void saddstep(QU &px, QU &py, double rxp, double rxm,
                                double2 ag, double2 ac)
{
  const double lw = rxp + rxm, lx = rxm - rxp;
  px.qr = fma(ag.x, lw, px.qr);
  px.qi = fma(ag.y, lw, px.qi);
  if (!DERIV) {
    px.ur = fma(ac.x, lw, px.ur);
    px.ui = fma(ac.y, lw, px.ui);
    py.qr = fma(-ac.y, lx, py.qr);
    py.qi = fma(ac.x, lx, py.qi);
  }
  py.ur = fma(ag.y, lx, py.ur);
  py.ui = fma(-ag.x, lx, py.ui);
}

template <int TG, bool DERIV, bool CROW = false>
__device__ void synth_spin_tile(const LegArgs &a, int m, int sb)
{
  constexpr int NALM = DERIV ? 1 : 2;
  const int lane = threadIdx.x;
  if (m > a.mmax) return;
  const int lmax = a.lmax, sp = a.spin;
  const int slot = sb * 64 + lane;
  const bool valid = slot < a.nslots;
  const int s = valid ? slot : a.nslots - 1;
  const double cth = a.cth[s], sth = a.sth[s];
  const bool active = valid && (a.mlim[s] >= m);
  const SpinStart st = spin_start(m, sp);

  QU p1[TG], p2[TG];
#pragma unroll
  for (int t = 0; t < TG; ++t) {
    p1[t] = QU{0., 0., 0., 0.};
    p2[t] = QU{0., 0., 0., 0.};
  }

  if (wave_any(active) && st.mhi <= lmax) {
    typedef std::conditional_t<CROW, URow, const double *> Row;
    typedef std::conditional_t<CROW, URow2, const double2 *> Row2;
    const Row fxb = Row(a.tab + 3 * (tab_ofs(m, lmax) - m));
    auto fx = [&](int l) { return Fx{fxb[3 * l], fxb[3 * l + 1], fxb[3 * l + 2]}; };
    const int nc = a.ncomp_alm;
    const Row2 al = Row2(a.alm + ((tri_ofs(m, lmax) - m) * nc + a.t0 * NALM));
    auto ag = [&](int l, int t) { return al[(int64_t)l * nc + t * NALM]; };
    auto ac = [&](int l, int t) {
      return DERIV ? make_double2(0., 0.) : al[(int64_t)l * nc + t * NALM + 1];
    };

    double rec1p = 0., rec1m = 0., rec2p, rec2m;
    int scalep, scalem;
    spin_init(cth, sth, sp, st, a.prefac[m], a.fscale[m], rec2p, rec2m, scalep,
              scalem);
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
      bool done = false;
      while (wave_any(scalep < 0 || scalem < 0)) {
#pragma unroll
        for (int t = 0; t < TG; ++t)
          saddstep<DERIV>(p1[t], p2[t], rec2p * cfp, rec2m * cfm, ag(l, t), ac(l, t));
        if (++l > lmax) { done = true; break; }
        rec_step(rec1p, rec1m, rec2p, rec2m, cth, fx(l));
#pragma unroll
        for (int t = 0; t < TG; ++t)
          saddstep<DERIV>(p2[t], p1[t], rec1p * cfp, rec1m * cfm, ag(l, t), ac(l, t));
        if (++l > lmax) { done = true; break; }
        rec_step(rec2p, rec2m, rec1p, rec1m, cth, fx(l));
        const bool a1 = rescale1(rec1p, rec2p, scalep);
        const bool a2 = rescale1(rec1m, rec2m, scalem);
        if (a1 || a2) { cfp = corfac(scalep); cfm = corfac(scalem); }
      }
      if (!done) {
        rec1p *= cfp; rec2p *= cfp; rec1m *= cfm; rec2m *= cfm;
        // hot loop, the next pair's loads requested after the first use of
        // the current ones (as in the spin-0 kernel); rows up to lmax+2 and
        // coefficients up to lmax+3 are readable
        Fx f1 = fx(l + 1), f2 = fx(l + 2);
        double2 g0[TG], c0[TG], g1[TG], c1[TG];
#pragma unroll
        for (int t = 0; t < TG; ++t) {
          g0[t] = ag(l, t); c0[t] = ac(l, t);
          g1[t] = ag(l + 1, t); c1[t] = ac(l + 1, t);
        }
        while (l < lmax) {
          rec_step(rec1p, rec1m, rec2p, rec2m, cth, f1);
          __builtin_amdgcn_sched_barrier(0);
          const Fx nf1 = fx(l + 3), nf2 = fx(l + 4);
          double2 ng0[TG], nc0[TG], ng1[TG], nc1[TG];
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            ng0[t] = ag(l + 2, t); nc0[t] = ac(l + 2, t);
            ng1[t] = ag(l + 3, t); nc1[t] = ac(l + 3, t);
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            saddstep<DERIV>(p1[t], p2[t], rec2p, rec2m, g0[t], c0[t]);
            saddstep<DERIV>(p2[t], p1[t], rec1p, rec1m, g1[t], c1[t]);
          }
          rec_step(rec2p, rec2m, rec1p, rec1m, cth, f2);
          f1 = nf1;
          f2 = nf2;
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            g0[t] = ng0[t]; c0[t] = nc0[t]; g1[t] = ng1[t]; c1[t] = nc1[t];
          }
          l += 2;
        }
        if (l == lmax) {
#pragma unroll
          for (int t = 0; t < TG; ++t)
            saddstep<DERIV>(p1[t], p2[t], rec2p, rec2m, g0[t], c0[t]);
        }
      }
    }
  }
  if (!valid) return;
  const bool pr = a.ispair[slot] != 0;
  const double sg = ((st.mhi - m + sp) & 1) ? -1. : 1.;
#pragma unroll
  for (int t = 0; t < TG; ++t) {
    const int cq = 2 * (a.t0 + t), cu = cq + 1;
    double2 qn = make_double2(0., 0.), un = qn, qs = qn, us = qn;
    if (active) {
      qn = make_double2(p1[t].qr + p2[t].qr, p1[t].qi + p2[t].qi);
      un = make_double2(p1[t].ur + p2[t].ur, p1[t].ui + p2[t].ui);
      qs = make_double2(sg * (p1[t].qr - p2[t].qr), sg * (p1[t].qi - p2[t].qi));
      us = make_double2(sg * (p1[t].ur - p2[t].ur), sg * (p1[t].ui - p2[t].ui));
    }
    a.phase[phase_index(cq, m, slot, 0, a.nslots, a.nmb)] = qn;
    a.phase[phase_index(cu, m, slot, 0, a.nslots, a.nmb)] = un;
    if (pr) {
      a.phase[phase_index(cq, m, slot, 1, a.nslots, a.nmb)] = qs;
      a.phase[phase_index(cu, m, slot, 1, a.nslots, a.nmb)] = us;
    }
  }
}

// ===========================================================================
// synthesis on the FP64 matrix cores
// ===========================================================================
// The wavefront layout of the mixed-precision syntheses: lane i of lane row r
// runs the recurrence of ring pair 16r+i over l in registers; each block of
// l values is turned (row_transpose4d) into the A operands of
// v_mfma_f64_16x16x4f64 (16 ring pairs x 4 l, one per lane row r), the a_lm
// rows of those l are its B operand (4 l x 16 columns), and the products
// accumulate in FP64 over all l.  The matrix cores take the contraction
// while the VALU runs the recurrence, for up to 32 columns per launch (16
// spin-0 or 8 spin-s transforms), so the recurrence runs once for all of
// them instead of once per group of 2 or 4.
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma64(double a, double b, d4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// On entry v[k] of lane 16r+i belongs to lane row r; on exit v[r] of lane
// 16k+i holds what was v[k] of lane 16r+i (the 4x4 transpose of lane rows
// against registers, on both halves of each double).
__device__ inline void row_transpose4d(double v[4])
{
  unsigned lo[4], hi[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v[k]);
    lo[k] = (unsigned)u;
    hi[k] = (unsigned)(u >> 32);
  }
  auto tr = [](unsigned (&w)[4]) {
    const auto a = __builtin_amdgcn_permlane32_swap(w[0], w[2], false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(w[1], w[3], false, false);
    const auto c = __builtin_amdgcn_permlane16_swap(a[0], b[0], false, false);
    const auto d = __builtin_amdgcn_permlane16_swap(a[1], b[1], false, false);
    w[0] = c[0];
    w[1] = c[1];
    w[2] = d[0];
    w[3] = d[1];
  };
  tr(lo);
  tr(hi);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    v[k] = __builtin_bit_cast(double, ((unsigned long long)hi[k] << 32) | lo[k]);
}

// stores N = t1+t2 and S = sg*(t1-t2) of the 64 ring pairs of a wavefront:
// lane n+16q holds rows 16r+4v+q (v = 0..3) of column n of accumulator r
template <int NCB>
__device__ inline void store_phases64(const LegArgs &a, int m, int sb, int comp0, int ncol,
                                      double sg, const d4 (&t1)[NCB][4],
                                      const d4 (&t2)[NCB][4])
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
        const int ring = sb * 64 + 16 * r + 4 * v + q;
        if (ring >= a.nslots) continue;
        const double x1 = t1[cb][r][v], x2 = t2[cb][r][v];
        ph[2 * phase_index(comp, m, ring, 0, a.nslots, a.nmb) + part] = x1 + x2;
        if (a.ispair[ring])
          ph[2 * phase_index(comp, m, ring, 1, a.nslots, a.nmb) + part] = sg * (x1 - x2);
      }
  }
}

// Spin 0, blocks of 8 l: accumulator 1 takes the even rows l0+2k with the
// a_lm rows l0+2k, accumulator 2 the odd ones (k = 0..3 the lane row); the
// products of a block are issued during the recurrence of the next one.
template <int NCB, bool CROW = false>
__device__ void synth64_spin0_tile(const LegArgs &a, int m, int sb)
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

  d4 acc1[NCB][4], acc2[NCB][4];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc1[cb][r] = acc2[cb][r] = d4{0., 0., 0., 0.};

  if (wave_any(active)) {
    typedef std::conditional_t<CROW, URow2, const double2 *> Row2;
    const Row2 rf = Row2(reinterpret_cast<const double2 *>(a.tab) + (tab_ofs(m, lmax) - m));
    const int lpad = lmax + kTabPad;   // last readable entry of the row
    // canonical a_lm row l of this m, first column of this launch; rows past
    // lmax (up to lmax+15) are the next m's or the slack rows, masked
    const double *arow = reinterpret_cast<const double *>(a.alm) +
                         ((int64_t)(tri_ofs(m, lmax) - m) * nc + a.t0) * 2;
    bool colok[NCB];
    int boff1[NCB], boff2[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int col = cb * 16 + n;
      colok[cb] = col < ncol;
      boff1[cb] = 2 * kq * 2 * nc + (colok[cb] ? col : 0);
      boff2[cb] = boff1[cb] + 2 * nc;
    }
    auto ldB = [&](int l0, double (&b1)[NCB], double (&b2)[NCB]) {
      const double *row = arow + (int64_t)l0 * 2 * nc;
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
    // below the IEEE range on every lane: advance without accumulating
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
      double2 rc[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) rc[k] = rf[min(l + k, lpad)];
      double bc1[NCB], bc2[NCB];
      ldB(l, bc1, bc2);
      double pv1[4] = {0., 0., 0., 0.}, pv2[4] = {0., 0., 0., 0.};
      double pb1[NCB], pb2[NCB];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) pb1[cb] = pb2[cb] = 0.;
      auto mfmas = [&]() {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            acc1[cb][r] = mfma64(pv1[r], pb1[cb], acc1[cb][r]);
            acc2[cb][r] = mfma64(pv2[r], pb2[cb], acc2[cb][r]);
          }
      };
      while (l <= lmax) {
        const int lb = l;
        double bn1[NCB], bn2[NCB];
        ldB(lb + 8, bn1, bn2);
        double2 rn[8];
        double v1[4], v2[4];
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
          v1[q] = o0;
          v2[q] = o1;
          l += 2;
          if (q == 0) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 8; ++k) rn[k] = rf[min(lb + 8 + k, lpad)];
            __builtin_amdgcn_sched_barrier(0);
          }
          if (q == 1) mfmas();   // the previous block's products
        }
        row_transpose4d(v1);
        row_transpose4d(v2);
        {
          const int l1 = lb + 2 * kq, l2 = l1 + 1;
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb) {
            pb1[cb] = (colok[cb] && l1 <= lmax) ? bc1[cb] : 0.;
            pb2[cb] = (colok[cb] && l2 <= lmax) ? bc2[cb] : 0.;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) { pv1[q] = v1[q]; pv2[q] = v2[q]; }
        if (scaled && wave_all(scale >= 0)) {
          lam1 *= cf; lam2 *= cf; cf = 1.;
          scaled = false;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) rc[k] = rn[k];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) { bc1[cb] = bn1[cb]; bc2[cb] = bn2[cb]; }
      }
      mfmas();
    }
  }
  store_phases64<NCB>(a, m, sb, a.t0, ncol, 1., acc1, acc2);
}

// Spin s (and ALM2MAP_DERIV1), blocks of 4 l: accumulator 1 takes the rows
// (w_l, x_l+1, w_l+2, x_l+3) with the a_lm rows (X_l, Y_l+1, X_l+2, Y_l+3),
// accumulator 2 the other way round, where lambda_w = lambda_+ + lambda_-,
// lambda_x = lambda_- - lambda_+, X = (E.re, E.im, B.re, B.im) and
// Y = (-B.im, B.re, E.im, -E.re) per Q/U column (libsharp/sharp_core_inc2.c:
// 282-306); N = t1+t2 and S = sg*(t1-t2) as in the scalar kernel.
template <int NCB, bool DERIV, bool CROW = false>
__device__ void synth64_spin_tile(const LegArgs &a, int m, int sb)
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

  d4 acc1[NCB][4], acc2[NCB][4];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc1[cb][r] = acc2[cb][r] = d4{0., 0., 0., 0.};

  if (wave_any(active) && st.mhi <= lmax) {
    typedef std::conditional_t<CROW, URow, const double *> Row;
    const Row fxb = Row(a.tab + 3 * (tab_ofs(m, lmax) - m));
    const int lpad = lmax + kTabPad;
    auto fx = [&](int l) {
      const int k = min(l, lpad);
      return Fx{fxb[3 * k], fxb[3 * k + 1], fxb[3 * k + 2]};
    };
    constexpr int NALM = DERIV ? 1 : 2;
    const double *arow = reinterpret_cast<const double *>(a.alm) +
                         ((int64_t)(tri_ofs(m, lmax) - m) * nc + a.t0 * NALM) * 2;
    // column col = 4t+q is (Q.re, Q.im, U.re, U.im)[q] of transform t: its
    // X and Y entries as offsets into the row with signs (offset 0, sign 0
    // for unused columns)
    int xo[NCB], yo[NCB];
    double xs[NCB], ys[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int col = cb * 16 + n, t = col >> 2, q = col & 3;
      const bool ok = col < ncol;
      if (!DERIV) {
        xo[cb] = ok ? col : 0;
        xs[cb] = ok ? 1. : 0.;
        yo[cb] = ok ? (col ^ 3) : 0;
        ys[cb] = ok ? ((q == 0 || q == 3) ? -1. : 1.) : 0.;
      } else {
        // no curl a_lm: X = (E.re, E.im, 0, 0), Y = (0, 0, E.im, -E.re)
        xo[cb] = ok ? 2 * t + (q & 1) : 0;
        xs[cb] = (ok && q < 2) ? 1. : 0.;
        yo[cb] = ok ? 2 * t + (q == 2 ? 1 : 0) : 0;
        ys[cb] = ok ? (q == 2 ? 1. : (q == 3 ? -1. : 0.)) : 0.;
      }
      // lane row kq reads a_lm row l0+kq
      xo[cb] += kq * 2 * nc;
      yo[cb] += kq * 2 * nc;
    }
    auto ldB = [&](int l0, double (&bx)[NCB], double (&by)[NCB]) {
      const double *row = arow + (int64_t)l0 * 2 * nc;
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        bx[cb] = row[xo[cb]];
        by[cb] = row[yo[cb]];
      }
    };

    double rec1p = 0., rec1m = 0., rec2p, rec2m;
    int scalep, scalem;
    spin_init(cth, sth, sp, st, a.prefac[m], a.fscale[m], rec2p, rec2m, scalep, scalem);
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
      Fx fc[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) fc[k] = fx(l + 1 + k);
      double bcx[NCB], bcy[NCB];
      ldB(l, bcx, bcy);
      const bool odd = (kq & 1) != 0;
      double pv1[4] = {0., 0., 0., 0.}, pv2[4] = {0., 0., 0., 0.};
      double pb1[NCB], pb2[NCB];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) pb1[cb] = pb2[cb] = 0.;
      auto mfmas = [&]() {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            acc1[cb][r] = mfma64(pv1[r], pb1[cb], acc1[cb][r]);
            acc2[cb][r] = mfma64(pv2[r], pb2[cb], acc2[cb][r]);
          }
      };
      while (l <= lmax) {
        const int lb = l;
        double bnx[NCB], bny[NCB];
        ldB(lb + 4, bnx, bny);
        Fx fn[4];
        double v1[4], v2[4];
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
          v1[2 * h] = lw0;
          v1[2 * h + 1] = lx1;
          v2[2 * h] = lx0;
          v2[2 * h + 1] = lw1;
          l += 2;
          if (h == 0) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) fn[k] = fx(lb + 5 + k);
            __builtin_amdgcn_sched_barrier(0);
            mfmas();   // the previous block's products
          }
        }
        row_transpose4d(v1);
        row_transpose4d(v2);
        {
          const bool okr = lb + kq <= lmax;
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb) {
            const double bx = okr ? xs[cb] * bcx[cb] : 0.;
            const double by = okr ? ys[cb] * bcy[cb] : 0.;
            pb1[cb] = odd ? by : bx;
            pb2[cb] = odd ? bx : by;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) { pv1[q] = v1[q]; pv2[q] = v2[q]; }
        if (scaled && wave_all(scalep >= 0 && scalem >= 0)) {
          rec1p *= cfp; rec2p *= cfp; rec1m *= cfm; rec2m *= cfm;
          cfp = cfm = 1.;
          scaled = false;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) fc[k] = fn[k];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) { bcx[cb] = bnx[cb]; bcy[cb] = bny[cb]; }
      }
      mfmas();
    }
  }
  const double sg = ((st.mhi - m + sp) & 1) ? -1. : 1.;
  store_phases64<NCB>(a, m, sb, 2 * a.t0, ncol, sg, acc1, acc2);
}

// ===========================================================================
// synthesis kernels: one job (grid over its m-blocks and slot blocks, the
// slot blocks of one m-block, which read the same coefficient and a_lm rows,
// on one XCD), or a batch (persistent grid over a work list)
// ===========================================================================
__device__ inline int wave_y() { return __builtin_amdgcn_readfirstlane(threadIdx.y); }

template <int TG>
__global__ void __launch_bounds__(64 * MB) k_synth_spin0(LegArgs a, int nsb)
{
  int mb, sb;
  if (xcd_map(blockIdx.x, a.nmb, nsb, mb, sb)) synth_spin0_tile<TG>(a, mb * MB + wave_y(), sb);
}

template <int TG, bool DERIV>
__global__ void __launch_bounds__(64 * MB) k_synth_spin(LegArgs a, int nsb)
{
  int mb, sb;
  if (xcd_map(blockIdx.x, a.nmb, nsb, mb, sb))
    synth_spin_tile<TG, DERIV>(a, mb * MB + wave_y(), sb);
}

template <int NCB>
__global__ void __launch_bounds__(64 * MB) k_synth64_spin0(LegArgs a, int nsb)
{
  int mb, sb;
  if (xcd_map(blockIdx.x, a.nmb, nsb, mb, sb))
    synth64_spin0_tile<NCB>(a, mb * MB + wave_y(), sb);
}

template <int NCB, bool DERIV>
__global__ void __launch_bounds__(64 * MB) k_synth64_spin(LegArgs a, int nsb)
{
  int mb, sb;
  if (xcd_map(blockIdx.x, a.nmb, nsb, mb, sb))
    synth64_spin_tile<NCB, DERIV>(a, mb * MB + wave_y(), sb);
}

// items: (m-block of blockDim.y orders m, slot block) tiles, one order m per
// wavefront (so that a workgroup writes whole phase lines), all of TG
// transforms (one kernel per group size: a kernel holding every size needs
// the registers of the largest and keeps half as many wavefronts resident);
// FLAT: one item per workgroup, else persistent
template <int TG, bool FLAT>
__global__ void __launch_bounds__(64 * MB) kb_synth_spin0(const LegArgs *__restrict__ jobs,
                                                          const BatchItem *__restrict__ items,
                                                          const int *__restrict__ xofs,
                                                          int *__restrict__ ctr)
{
  auto tile = [&](const BatchItem &it) {
    synth_spin0_tile<TG, true>(job_args(jobs, it), it.major * blockDim.y + wave_y(), it.minor);
  };
  if constexpr (FLAT) {
    for_flat_item(items, xofs, tile);
  } else {
    __shared__ int fetched;
    for_group_items(items, xofs, ctr, fetched, tile);
  }
}

template <int TG, bool DERIV, bool FLAT>
__global__ void __launch_bounds__(64 * MB) kb_synth_spin(const LegArgs *__restrict__ jobs,
                                                         const BatchItem *__restrict__ items,
                                                         const int *__restrict__ xofs,
                                                         int *__restrict__ ctr)
{
  auto tile = [&](const BatchItem &it) {
    synth_spin_tile<TG, DERIV, true>(job_args(jobs, it), it.major * blockDim.y + wave_y(),
                                     it.minor);
  };
  if constexpr (FLAT) {
    for_flat_item(items, xofs, tile);
  } else {
    __shared__ int fetched;
    for_group_items(items, xofs, ctr, fetched, tile);
  }
}

// The matrix-core FP64 syntheses over a batch's items (the tiles of
// synth64_*_tile: NCB blocks of 16 real columns, any number of transforms up
// to them), coefficient and a_lm rows read through the constant address space
template <int NCB, bool FLAT>
__global__ void __launch_bounds__(64 * MB) kb_synth64_spin0(const LegArgs *__restrict__ jobs,
                                                            const BatchItem *__restrict__ items,
                                                            const int *__restrict__ xofs,
                                                            int *__restrict__ ctr)
{
  auto tile = [&](const BatchItem &it) {
    synth64_spin0_tile<NCB, true>(job_args(jobs, it), it.major * blockDim.y + wave_y(), it.minor);
  };
  if constexpr (FLAT) {
    for_flat_item(items, xofs, tile);
  } else {
    __shared__ int fetched;
    for_group_items(items, xofs, ctr, fetched, tile);
  }
}

template <int NCB, bool DERIV, bool FLAT>
__global__ void __launch_bounds__(64 * MB) kb_synth64_spin(const LegArgs *__restrict__ jobs,
                                                           const BatchItem *__restrict__ items,
                                                           const int *__restrict__ xofs,
                                                           int *__restrict__ ctr)
{
  auto tile = [&](const BatchItem &it) {
    synth64_spin_tile<NCB, DERIV, true>(job_args(jobs, it), it.major * blockDim.y + wave_y(),
                                        it.minor);
  };
  if constexpr (FLAT) {
    for_flat_item(items, xofs, tile);
  } else {
    __shared__ int fetched;
    for_group_items(items, xofs, ctr, fetched, tile);
  }
}

// ===========================================================================
// spin 0 analysis
// ===========================================================================
constexpr int kChunk = 8;   // l values per reduction chunk

template <int TG>
__global__ void __launch_bounds__(256) k_anal_spin0(LegArgs a)
{
  constexpr int K = kChunk;
  constexpr int NV = 2 * K * TG;          // doubles per lane per chunk
  constexpr int BMASK = 64 / NV - 1;      // low lane bits of replicated results
  static_assert(NV <= 64, "too many values per chunk");
  __shared__ double red[2][4][NV];

  const int m = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lmax = a.lmax, nc = a.ncomp_alm;
  const int nchunks = (lmax - m + K) / K;
  const int ngroups = (a.nslots + 255) / 256;
  double *out = reinterpret_cast<double *>(a.almout + ((tri_ofs(m, lmax) - m) * nc + a.t0));
  const double2 *rf = reinterpret_cast<const double2 *>(a.tab) + (tab_ofs(m, lmax) - m);

  // zero the output row of this m (all components of this launch)
  for (int i = threadIdx.x; i < (lmax - m + 1) * TG * 2; i += 256) {
    const int l = m + i / (2 * TG), r = i % (2 * TG);
    out[(int64_t)l * nc * 2 + r] = 0.;
  }
  __syncthreads();

  const bool writer = (w == 0) && (lane < NV);
  const int wt = lane / (2 * K), wj = lane % (2 * K);
  const int myidx = xred_index<NV>(lane);
  const bool holder = (lane & BMASK) == 0;
  int bufsel = 0;

  for (int g = 0; g < ngroups; ++g) {
    const int slot = g * 256 + w * 64 + lane;
    const bool valid = slot < a.nslots;
    const int s = valid ? slot : a.nslots - 1;
    const bool active = valid && (a.mlim[s] >= m);
    if (!__syncthreads_or(active)) continue;
    const double cth = a.cth[s], sth = a.sth[s];
    double p1r[TG], p1i[TG], p2r[TG], p2i[TG];
    const bool pr = a.ispair[s] != 0;
#pragma unroll
    for (int t = 0; t < TG; ++t) {
      p1r[t] = p1i[t] = p2r[t] = p2i[t] = 0.;
      if (active) {
        const double2 n = a.phase[phase_index(a.t0 + t, m, slot, 0, a.nslots, a.nmb)];
        double2 sv = make_double2(0., 0.);
        if (pr) sv = a.phase[phase_index(a.t0 + t, m, slot, 1, a.nslots, a.nmb)];
        p1r[t] = n.x + sv.x; p1i[t] = n.y + sv.y;
        p2r[t] = n.x - sv.x; p2i[t] = n.y - sv.y;
      }
    }
    double lam1 = 0., lam2;
    int scale;
    mypow(sth, m, lam2, scale);
    const double mf = a.mfac[m];
    lam2 *= (m & 1) ? -mf : mf;
    normalize(lam2, scale, kFTol);
    // 0: advancing below the IEEE range, 1: scaled, 2: unscaled, 3: idle.
    // As in libsharp the advance phase only runs while every lane is below
    // the IEEE range (libsharp/sharp_core_inc.c:192-205).
    int mode = !wave_any(active) ? 3 : (wave_all(scale < 1) ? 0 : 1);
    double cf = corfac(scale);
    int l = m;

    for (int c = 0; c < nchunks; ++c) {
      double acc[NV];
#pragma unroll
      for (int i = 0; i < NV; ++i) acc[i] = 0.;
      bool contrib = false;
      if (mode == 2) {
        contrib = true;
#pragma unroll
        for (int k = 0; k < K; k += 2) {
          const double2 r0 = rf[l], r1 = rf[l + 1];
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            acc[t * 2 * K + 2 * k] = lam2 * p1r[t];
            acc[t * 2 * K + 2 * k + 1] = lam2 * p1i[t];
          }
          lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
#pragma unroll
          for (int t = 0; t < TG; ++t) {
            acc[t * 2 * K + 2 * k + 2] = lam1 * p2r[t];
            acc[t * 2 * K + 2 * k + 3] = lam1 * p2i[t];
          }
          lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
          l += 2;
        }
      } else if (mode < 2) {
#pragma unroll
        for (int k = 0; k < K; k += 2) {
          const double2 r0 = rf[l], r1 = rf[l + 1];
          if (mode == 0) {
            lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
            lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
            if (fabs(lam2) > kFTol) { lam1 *= kFSmall; lam2 *= kFSmall; ++scale; }
            if (wave_any(scale >= 1)) { mode = 1; cf = corfac(scale); }
          } else {
            contrib = true;
            const double t2 = lam2 * cf;
#pragma unroll
            for (int t = 0; t < TG; ++t) {
              acc[t * 2 * K + 2 * k] = t2 * p1r[t];
              acc[t * 2 * K + 2 * k + 1] = t2 * p1i[t];
            }
            lam1 = (cth * lam2) * r0.x - lam1 * r0.y;
            const double t1 = lam1 * cf;
#pragma unroll
            for (int t = 0; t < TG; ++t) {
              acc[t * 2 * K + 2 * k + 2] = t1 * p2r[t];
              acc[t * 2 * K + 2 * k + 3] = t1 * p2i[t];
            }
            lam2 = (cth * lam1) * r1.x - lam2 * r1.y;
            if (fabs(lam2) > kFTol) {
              lam1 *= kFSmall; lam2 *= kFSmall; ++scale;
              cf = corfac(scale);
            }
          }
          l += 2;
        }
        if (mode == 1 && wave_all(scale >= 0)) {
          lam1 *= cf; lam2 *= cf;
          mode = 2;
        }
      }
      if (contrib) XRed<NV, 32>::run(acc, lane);
      if (holder) red[bufsel][w][myidx] = contrib ? acc[0] : 0.;
      __syncthreads();
      if (writer) {
        const double v = red[bufsel][0][lane] + red[bufsel][1][lane] +
                         red[bufsel][2][lane] + red[bufsel][3][lane];
        const int ll = m + K * c + wj / 2;
        if (ll <= lmax) out[((int64_t)ll * nc + wt) * 2 + (wj & 1)] += v;
      }
      bufsel ^= 1;
    }
  }
}

// ===========================================================================
// spin s analysis
// ===========================================================================
template <int TG>
__global__ void __launch_bounds__(256) k_anal_spin(LegArgs a)
{
  constexpr int K = kChunk;
  constexpr int NV = 4 * K * TG;
  constexpr int BMASK = 64 / NV - 1;
  static_assert(NV <= 64, "too many values per chunk");
  __shared__ double red[2][4][NV];

  const int m = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lmax = a.lmax, nc = a.ncomp_alm, sp = a.spin;
  const SpinStart st = spin_start(m, sp);
  const int nchunks = (st.mhi <= lmax) ? (lmax - st.mhi + K) / K : 0;
  const int ngroups = (a.nslots + 255) / 256;
  double *out = reinterpret_cast<double *>(a.almout + ((tri_ofs(m, lmax) - m) * nc + 2 * a.t0));
  const double *fxb = a.tab + 3 * (tab_ofs(m, lmax) - m);
  auto fx = [&](int l) { return Fx{fxb[3 * l], fxb[3 * l + 1], fxb[3 * l + 2]}; };

  for (int i = threadIdx.x; i < (lmax - m + 1) * TG * 4; i += 256) {
    const int l = m + i / (4 * TG), r = i % (4 * TG);
    out[(int64_t)l * nc * 2 + r] = 0.;
  }
  __syncthreads();

  const bool writer = (w == 0) && (lane < NV);
  const int wt = lane / (4 * K), wj = lane % (4 * K);
  const int myidx = xred_index<NV>(lane);
  const bool holder = (lane & BMASK) == 0;
  const double sg = ((st.mhi - m + sp) & 1) ? -1. : 1.;
  int bufsel = 0;

  for (int g = 0; g < ngroups && nchunks > 0; ++g) {
    const int slot = g * 256 + w * 64 + lane;
    const bool valid = slot < a.nslots;
    const int s = valid ? slot : a.nslots - 1;
    const bool active = valid && (a.mlim[s] >= m);
    if (!__syncthreads_or(active)) continue;
    const double cth = a.cth[s], sth = a.sth[s];
    const bool pr = a.ispair[s] != 0;
    // P[0] = N+S (used for l-mhi even), P[1] = N-S
    QU P[2][TG];
#pragma unroll
    for (int t = 0; t < TG; ++t) {
      P[0][t] = QU{0., 0., 0., 0.};
      P[1][t] = QU{0., 0., 0., 0.};
      if (active) {
        const int cq = 2 * (a.t0 + t), cu = cq + 1;
        const double2 nq = a.phase[phase_index(cq, m, slot, 0, a.nslots, a.nmb)];
        const double2 nu = a.phase[phase_index(cu, m, slot, 0, a.nslots, a.nmb)];
        double2 sq = make_double2(0., 0.), su = sq;
        if (pr) {
          sq = a.phase[phase_index(cq, m, slot, 1, a.nslots, a.nmb)];
          su = a.phase[phase_index(cu, m, slot, 1, a.nslots, a.nmb)];
          sq.x *= sg; sq.y *= sg; su.x *= sg; su.y *= sg;
        }
        P[0][t] = QU{nq.x + sq.x, nq.y + sq.y, nu.x + su.x, nu.y + su.y};
        P[1][t] = QU{nq.x - sq.x, nq.y - sq.y, nu.x - su.x, nu.y - su.y};
      }
    }
    double rec1p = 0., rec1m = 0., rec2p, rec2m;
    int scalep, scalem;
    spin_init(cth, sth, sp, st, a.prefac[m], a.fscale[m], rec2p, rec2m, scalep,
              scalem);
    int mode = !wave_any(active) ? 3 : (wave_all(scalep < 1 && scalem < 1) ? 0 : 1);
    double cfp = corfac(scalep), cfm = corfac(scalem);
    int l = st.mhi;

    // a(E,B) += saddstep2(px, py, rp, rm) (libsharp/sharp_core_inc2.c:339-364)
    auto accum = [&](double *acc, int k, int x, double rp, double rm) {
      const double lw = rp + rm, lx = rm - rp;
#pragma unroll
      for (int t = 0; t < TG; ++t) {
        const QU &px = P[x][t], &py = P[x ^ 1][t];
        double *o = acc + t * 4 * K + 4 * k;
        o[0] = px.qr * lw - py.ui * lx;
        o[1] = px.qi * lw + py.ur * lx;
        o[2] = px.ur * lw + py.qi * lx;
        o[3] = px.ui * lw - py.qr * lx;
      }
    };

    for (int c = 0; c < nchunks; ++c) {
      double acc[NV];
#pragma unroll
      for (int i = 0; i < NV; ++i) acc[i] = 0.;
      bool contrib = false;
      if (mode == 2) {
        contrib = true;
#pragma unroll
        for (int k = 0; k < K; k += 2) {
          accum(acc, k, 0, rec2p, rec2m);
          rec_step(rec1p, rec1m, rec2p, rec2m, cth, fx(l + 1));
          accum(acc, k + 1, 1, rec1p, rec1m);
          rec_step(rec2p, rec2m, rec1p, rec1m, cth, fx(l + 2));
          l += 2;
        }
      } else if (mode < 2) {
#pragma unroll
        for (int k = 0; k < K; k += 2) {
          if (mode == 0) {
            rec_step(rec1p, rec1m, rec2p, rec2m, cth, fx(l + 1));
            rec_step(rec2p, rec2m, rec1p, rec1m, cth, fx(l + 2));
            rescale1(rec1p, rec2p, scalep);
            rescale1(rec1m, rec2m, scalem);
            if (wave_any(scalep >= 1 || scalem >= 1)) {
              mode = 1;
              cfp = corfac(scalep); cfm = corfac(scalem);
            }
          } else {
            contrib = true;
            accum(acc, k, 0, rec2p * cfp, rec2m * cfm);
            rec_step(rec1p, rec1m, rec2p, rec2m, cth, fx(l + 1));
            accum(acc, k + 1, 1, rec1p * cfp, rec1m * cfm);
            rec_step(rec2p, rec2m, rec1p, rec1m, cth, fx(l + 2));
            const bool a1 = rescale1(rec1p, rec2p, scalep);
            const bool a2 = rescale1(rec1m, rec2m, scalem);
            if (a1 || a2) { cfp = corfac(scalep); cfm = corfac(scalem); }
          }
          l += 2;
        }
        if (mode == 1 && wave_all(scalep >= 0 && scalem >= 0)) {
          rec1p *= cfp; rec2p *= cfp; rec1m *= cfm; rec2m *= cfm;
          mode = 2;
        }
      }
      if (contrib) XRed<NV, 32>::run(acc, lane);
      if (holder) red[bufsel][w][myidx] = contrib ? acc[0] : 0.;
      __syncthreads();
      if (writer) {
        const double v = red[bufsel][0][lane] + red[bufsel][1][lane] +
                         red[bufsel][2][lane] + red[bufsel][3][lane];
        const int ll = st.mhi + K * c + wj / 4;
        const int q = wj % 4;   // 0: E re, 1: E im, 2: B re, 3: B im
        if (ll <= lmax)
          out[((int64_t)ll * nc + 2 * wt + (q >> 1)) * 2 + (q & 1)] += v;
      }
      bufsel ^= 1;
    }
  }
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

void launch_synth(const LegArgs &a, int ntg, hipStream_t st)
{
  // the slot blocks of one m block (which read the same coefficient and a_lm
  // rows) on one XCD
  const int nsb = (a.nslots + 63) / 64;
  const dim3 grid(xcd_grid(a.nmb, nsb));
  const dim3 block(64, MB);
  if (a.spin == 0) {
    switch (ntg) {
    case 1: hipLaunchKernelGGL(k_synth_spin0<1>, grid, block, 0, st, a, nsb); break;
    case 2: hipLaunchKernelGGL(k_synth_spin0<2>, grid, block, 0, st, a, nsb); break;
    case 4: hipLaunchKernelGGL(k_synth_spin0<4>, grid, block, 0, st, a, nsb); break;
    default: SHP_FAIL("bad transform group");
    }
  } else if (a.deriv1) {
    switch (ntg) {
    case 1: hipLaunchKernelGGL((k_synth_spin<1, true>), grid, block, 0, st, a, nsb); break;
    case 2: hipLaunchKernelGGL((k_synth_spin<2, true>), grid, block, 0, st, a, nsb); break;
    default: SHP_FAIL("bad transform group");
    }
  } else {
    switch (ntg) {
    case 1: hipLaunchKernelGGL((k_synth_spin<1, false>), grid, block, 0, st, a, nsb); break;
    case 2: hipLaunchKernelGGL((k_synth_spin<2, false>), grid, block, 0, st, a, nsb); break;
    default: SHP_FAIL("bad transform group");
    }
  }
  check_launch("synthesis");
}

void launch_synth64(const LegArgs &a, hipStream_t st)
{
  const int ncol = (a.spin == 0 ? 2 : 4) * a.ntl;
  if (a.ntl < 1 || ncol > 32) SHP_FAIL("bad transform group");
  const int nsb = (a.nslots + 63) / 64;
  const dim3 grid(xcd_grid(a.nmb, nsb));
  const dim3 block(64, MB);
  const bool two = ncol > 16;
  if (a.spin == 0) {
    if (two) hipLaunchKernelGGL(k_synth64_spin0<2>, grid, block, 0, st, a, nsb);
    else hipLaunchKernelGGL(k_synth64_spin0<1>, grid, block, 0, st, a, nsb);
  } else if (a.deriv1) {
    if (two) hipLaunchKernelGGL((k_synth64_spin<2, true>), grid, block, 0, st, a, nsb);
    else hipLaunchKernelGGL((k_synth64_spin<1, true>), grid, block, 0, st, a, nsb);
  } else {
    if (two) hipLaunchKernelGGL((k_synth64_spin<2, false>), grid, block, 0, st, a, nsb);
    else hipLaunchKernelGGL((k_synth64_spin<1, false>), grid, block, 0, st, a, nsb);
  }
  check_launch("FP64 matrix-core synthesis");
}

template <bool FLAT>
static void batch_synth_f64(int kind, int tg, dim3 b, int grid, int flat_grid, hipStream_t st,
                            const LegArgs *jobs, const BatchItem *items, const int *xofs,
                            int *ctr)
{
  auto go = [&](auto kernel) {
    launch_batch_kernel(kernel, FLAT, b, grid, flat_grid, st, jobs, items, xofs, ctr);
  };
  if (kind == kBatchSynth0) {
    if (tg == 4) go(kb_synth_spin0<4, FLAT>);
    else if (tg == 2) go(kb_synth_spin0<2, FLAT>);
    else go(kb_synth_spin0<1, FLAT>);
  } else if (kind == kBatchSynthS) {
    if (tg == 2) go(kb_synth_spin<2, false, FLAT>); else go(kb_synth_spin<1, false, FLAT>);
  } else {
    if (tg == 2) go(kb_synth_spin<2, true, FLAT>); else go(kb_synth_spin<1, true, FLAT>);
  }
}

template <bool FLAT>
static void batch_synth64(int kind, int ncb, dim3 b, int grid, int flat_grid, hipStream_t st,
                          const LegArgs *jobs, const BatchItem *items, const int *xofs, int *ctr)
{
  auto go = [&](auto kernel) {
    launch_batch_kernel(kernel, FLAT, b, grid, flat_grid, st, jobs, items, xofs, ctr);
  };
  const bool two = ncb == 2;
  if (kind == kBatchSynth0) {
    if (two) go(kb_synth64_spin0<2, FLAT>); else go(kb_synth64_spin0<1, FLAT>);
  } else if (kind == kBatchSynthS) {
    if (two) go(kb_synth64_spin<2, false, FLAT>); else go(kb_synth64_spin<1, false, FLAT>);
  } else {
    if (two) go(kb_synth64_spin<2, true, FLAT>); else go(kb_synth64_spin<1, true, FLAT>);
  }
}

void launch_batch_synth_f64(int kind, int tg, int shape, bool flat, const LegArgs *jobs,
                            const BatchItem *items, const int *xofs, int *ctr, int grid,
                            int flat_grid, hipStream_t st)
{
  const dim3 b(64, shape);
  if (flat) batch_synth_f64<true>(kind, tg, b, grid, flat_grid, st, jobs, items, xofs, ctr);
  else batch_synth_f64<false>(kind, tg, b, grid, flat_grid, st, jobs, items, xofs, ctr);
  check_launch("batched FP64 synthesis");
}

void launch_batch_synth64(int kind, int ncb, int shape, bool flat, const LegArgs *jobs,
                          const BatchItem *items, const int *xofs, int *ctr, int grid,
                          int flat_grid, hipStream_t st)
{
  if (ncb != 1 && ncb != 2) SHP_FAIL("bad column block count");
  const dim3 b(64, shape);
  if (flat) batch_synth64<true>(kind, ncb, b, grid, flat_grid, st, jobs, items, xofs, ctr);
  else batch_synth64<false>(kind, ncb, b, grid, flat_grid, st, jobs, items, xofs, ctr);
  check_launch("batched FP64 matrix-core synthesis");
}

void launch_anal(const LegArgs &a, int ntg, hipStream_t st)
{
  const dim3 grid(a.mmax + 1);
  const dim3 block(256);
  if (a.spin == 0) {
    switch (ntg) {
    case 1: hipLaunchKernelGGL(k_anal_spin0<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_anal_spin0<2>, grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_anal_spin0<4>, grid, block, 0, st, a); break;
    default: SHP_FAIL("bad transform group");
    }
  } else {
    switch (ntg) {
    case 1: hipLaunchKernelGGL(k_anal_spin<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_anal_spin<2>, grid, block, 0, st, a); break;
    default: SHP_FAIL("bad transform group");
    }
  }
  check_launch("analysis");
}

} // namespace shp

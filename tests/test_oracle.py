"""Pins the C restatement (oracle/sht_oracle.c) on the CPU.

Each check is chosen so that a plausible mistake in the restatement (a sign,
an index, a missing term, a wrong normalisation) fails one of them:

* the reference's own golden values (libsharp/sharp_testsuite.c:337-443);
* the reference build itself (oracle/_ref) on random inputs, every job type
  and flag;
* closed forms: spherical harmonics from scipy and mpmath (deep in the
  exponent-scaled regime as well), spin-weighted harmonics from Goldberg's
  formula, numpy's FFT;
* exact properties: Gauss-Legendre quadrature makes analysis invert synthesis
  (libsharp/sharp_testsuite.c:490-529), and analysis without weights is the
  adjoint of synthesis.
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from sht_cases import flip_theta, healpix, iso, rect, run_case, tri

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sign_scale_goldens(oracle):
    gold = json.load(open(os.path.join(HERE, "golden", "sign_scale.json")))
    lmax, nr, pp = gold["lmax"], gold["nrings"], gold["ppring"]
    npix = nr * pp
    nalm = (lmax + 1) * (lmax + 2) // 2
    g = oracle.iso("gauss", nr, pp)
    flip_theta(g)
    a = oracle.triangular(lmax)
    for ntrans in (1, 2):
        for case in gold["cases"]:
            nmaps = len(case["maps"])
            na = 1 if (case["spin"] == 0 or case["type"] == O.DERIV1) else 2
            alms = [np.full(nalm, 1 + 1j) for _ in range(ntrans * na)]
            maps = [np.zeros(npix) for _ in range(ntrans * nmaps)]
            oracle.execute(case["type"], case["spin"], alms, maps, g, a, ntrans, O.DP)
            for t in range(ntrans):
                for k, want in enumerate(case["maps"]):
                    got = maps[t * nmaps + k][[0, npix // 2, npix - 1]]
                    np.testing.assert_allclose(got, want, rtol=max(case["rtol"]))


def test_reference_build_goldens(ref):
    """The reference compiled in place reproduces its own golden values."""
    gold = json.load(open(os.path.join(HERE, "golden", "sign_scale.json")))
    lmax, nr, pp = gold["lmax"], gold["nrings"], gold["ppring"]
    npix = nr * pp
    nalm = (lmax + 1) * (lmax + 2) // 2
    g = ref.iso("gauss", nr, pp)
    flip_theta(g)
    a = ref.triangular(lmax)
    for case in gold["cases"]:
        na = 1 if (case["spin"] == 0 or case["type"] == O.DERIV1) else 2
        alms = [np.full(nalm, 1 + 1j) for _ in range(na)]
        maps = [np.zeros(npix) for _ in range(len(case["maps"]))]
        ref.execute(case["type"], case["spin"], alms, maps, g, a, 1, O.DP)
        for k, want in enumerate(case["maps"]):
            np.testing.assert_allclose(maps[k][[0, npix // 2, npix - 1]], want,
                                       rtol=max(case["rtol"]))


def _cmp(checker, lib, *args, tol=1e-12, **kw):
    pa, pm = run_case(checker, *args, **kw)
    ca, cm = run_case(lib, *args, **kw)
    errs = [rel_l2(x, y) for x, y in zip(pa + pm, ca + cm)]
    assert max(errs) < tol, errs


@pytest.mark.parametrize("spin", [0, 1, 2, 3])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM, O.Yt, O.WY])
def test_oracle_vs_reference_healpix(oracle, ref, spin, jobtype):
    _cmp(ref, oracle, healpix(8), tri(20), jobtype, spin, 2, O.DP)


@pytest.mark.parametrize("extra", [0, O.ADD, O.REAL_HARMONICS])
def test_oracle_vs_reference_flags(oracle, ref, extra):
    for jt in (O.ALM2MAP, O.MAP2ALM):
        _cmp(ref, oracle, iso("gauss", 17, 33), tri(16), jt, 2, 1, O.DP | extra)
        _cmp(ref, oracle, healpix(4), rect(10, 6), jt, 0, 3, O.DP | extra)
    _cmp(ref, oracle, healpix(4), tri(12), O.DERIV1, 1, 1, O.DP | extra)
    _cmp(ref, oracle, healpix(4), tri(12), O.ALM2MAP, 0, 1, extra, tol=1e-5)
    _cmp(ref, oracle, iso("gauss", 13, 13), tri(12), O.MAP2ALM, 2, 1,
         O.DP | O.NO_FFT | extra)


@pytest.mark.parametrize("kind", ["fejer1", "cc", "fejer2", "mw"])
def test_oracle_vs_reference_grids(oracle, ref, kind):
    for jt in (O.ALM2MAP, O.MAP2ALM):
        _cmp(ref, oracle, iso(kind, 21, 30), tri(12), jt, 0, 1, O.DP)


def test_oracle_vs_reference_spin30(oracle, ref):
    for jt in (O.ALM2MAP, O.MAP2ALM):
        _cmp(ref, oracle, iso("gauss", 48, 95), tri(47), jt, 30, 1, O.DP)


# --- closed forms -------------------------------------------------------------
def test_ylm_spin0_vs_scipy(oracle):
    """lambda_lm(theta) = Y_lm(theta, 0), Condon-Shortley phase included."""
    from scipy.special import sph_harm_y
    lmax = 40
    for theta in (0.05, 0.7, 1.3, math.pi / 2, 2.4, 3.1):
        for m in (0, 1, 5, 17, 40):
            lam = O.ylm_spin0(oracle, lmax, m, math.cos(theta), math.sin(theta))
            ls = np.arange(m, lmax + 1)
            want = np.real(sph_harm_y(ls, m, theta, 0.0))
            np.testing.assert_allclose(lam, want, rtol=1e-11, atol=1e-14)


def _ylm_mpmath(l, m, theta):
    """Orthonormal Y_lm(theta, 0) with Condon-Shortley phase, 40 digits."""
    mp.mp.dps = 40
    x = mp.cos(mp.mpf(theta))
    norm = mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.factorial(l - m) / mp.factorial(l + m))
    return norm * mp.legenp(l, m, x)   # legenp carries (-1)^m


@pytest.mark.parametrize("l,m,theta", [(600, 500, 0.9), (1000, 700, 0.75),
                                       (2000, 1900, 1.2), (400, 300, 0.8)])
def test_ylm_spin0_scaled_regime(oracle, l, m, theta):
    """Values far below 2^-1022 at the start of the recurrence: the exponent
    scaling must carry them to the right magnitude."""
    lam = O.ylm_spin0(oracle, l, m, math.cos(theta), math.sin(theta))
    want = float(_ylm_mpmath(l, m, theta))
    got = lam[l - m]
    if abs(want) < 2.0 ** -60:
        assert abs(got) < 2.0 ** -59   # below the range libsharp evaluates
    else:
        assert abs(got - want) <= 1e-10 * abs(want), (got, want)


def _sylm_goldberg(s, l, m, theta):
    """Spin-weighted sY_lm(theta, 0) by Goldberg et al. (1967), J. Math. Phys.
    8, 2155, eq. (3.1), evaluated in mpmath."""
    mp.mp.dps = 40
    th = mp.mpf(theta)
    pref = (-1) ** m * mp.sqrt(mp.factorial(l + m) * mp.factorial(l - m) * (2 * l + 1)
                               / (4 * mp.pi * mp.factorial(l + s) * mp.factorial(l - s)))
    tot = mp.mpf(0)
    for r in range(0, l - s + 1):
        k = r + s - m
        if k < 0 or k > l + s:
            continue
        tot += (mp.binomial(l - s, r) * mp.binomial(l + s, k) * (-1) ** (l - r - s)
                * mp.cot(th / 2) ** (2 * r + s - m))
    return pref * mp.sin(th / 2) ** (2 * l) * tot


def test_spin2_transform_vs_goldberg(oracle):
    """One E-mode coefficient through spin-2 synthesis equals the analytic
    combination of 2Y_lm and -2Y_lm: with a_lm^E = 1 (m>0, and its -m partner
    implied by reality) Q + iU = -sum_{+-m} a^E_lm 2Y_lm.  The map at phi=0 of
    a single (l, m) mode is compared ring by ring."""
    lmax, l, m = 10, 7, 3
    theta = np.array([0.3, 1.1, 1.9, 2.7])
    nph = 2 * lmax + 1
    nr = len(theta)
    g = oracle.general_geom(np.full(nr, nph), np.arange(nr) * nph, np.ones(nr),
                            np.zeros(nr), theta)
    a = oracle.triangular(lmax)
    nalm = (lmax + 1) * (lmax + 2) // 2
    idx = m * (2 * lmax + 1 - m) // 2 + l
    aE = np.zeros(nalm, np.complex128)
    aE[idx] = 1.0
    aB = np.zeros(nalm, np.complex128)
    maps = [np.zeros(nr * nph), np.zeros(nr * nph)]
    oracle.execute(O.ALM2MAP, 2, [aE, aB], maps, g, a, 1, O.DP)
    gi = g.contents
    for i in range(gi.npairs):
        for r in (gi.pair[i].r1, gi.pair[i].r2):
            if r.nph <= 0:
                continue
            th = r.theta
            q = maps[0][r.ofs]
            u = maps[1][r.ofs]
            # E-mode: Q + iU = -(a 2Y_lm + conj(a) (-1)^m 2Y_l-m) at phi=0
            y_p = complex(_sylm_goldberg(2, l, m, th))
            y_m = complex(_sylm_goldberg(2, l, -m, th))
            want = -(y_p + (-1) ** m * y_m)
            np.testing.assert_allclose([q, u], [want.real, want.imag], atol=1e-12)
    oracle.destroy_geom(g)
    oracle.destroy_alm(a)


def test_rdft_vs_numpy(oracle):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 4, 5, 12, 28, 97, 200, 257):
        x = rng.normal(size=n)
        f = O.rdft_forward(oracle, x)
        X = np.fft.rfft(x)
        want = np.zeros(n)
        want[0] = X[0].real
        for k in range(1, (n + 1) // 2):
            want[2 * k - 1] = X[k].real
            want[2 * k] = X[k].imag
        if n % 2 == 0 and n > 1:
            want[n - 1] = X[n // 2].real
        np.testing.assert_allclose(f, want, atol=1e-12 * n)
        back = O.rdft_backward(oracle, f)
        np.testing.assert_allclose(back / n, x, atol=3e-14 * n)


@pytest.mark.parametrize("spin", [0, 1, 2, 3, 30])
def test_gauss_roundtrip_exact(oracle, spin):
    """acctest criterion (libsharp/sharp_testsuite.c:490-529): analysis
    inverts synthesis on a Gauss-Legendre grid with nlat = lmax+1,
    nlon = 2 lmax + 1 to 1e-10."""
    lmax = 47
    g = oracle.iso("gauss", lmax + 1, 2 * lmax + 2)
    a = oracle.triangular(lmax)
    nc = 1 if spin == 0 else 2
    rng = np.random.default_rng(11)
    alms = O.random_alm(lmax, lmax, rng, spin, nc)
    maps = [np.zeros((lmax + 1) * (2 * lmax + 2)) for _ in range(nc)]
    oracle.execute(O.ALM2MAP, spin, [x.copy() for x in alms], maps, g, a, 1, O.DP)
    back = [np.zeros_like(x) for x in alms]
    oracle.execute(O.MAP2ALM, spin, back, maps, g, a, 1, O.DP)
    for x, y in zip(back, alms):
        assert rel_l2(x, y) < 1e-10
        assert np.abs(x - y).max() < 1e-10


@pytest.mark.parametrize("spin", [0, 2])
def test_adjoint_identity(oracle, spin):
    """<Y a, x> = sum_m c_m Re<a_m, (Y^T x)_m>, c_0 = 1, c_{m>0} = 2 (real
    maps, m >= 0 storage)."""
    lmax = 24
    g = oracle.healpix(8)
    a = oracle.triangular(lmax)
    nc = 1 if spin == 0 else 2
    rng = np.random.default_rng(5)
    alms = O.random_alm(lmax, lmax, rng, spin, nc)
    npix = 12 * 64
    ya = [np.zeros(npix) for _ in range(nc)]
    oracle.execute(O.ALM2MAP, spin, [x.copy() for x in alms], ya, g, a, 1, O.DP)
    x = [rng.normal(size=npix) for _ in range(nc)]
    ytx = [np.zeros_like(alms[0]) for _ in range(nc)]
    oracle.execute(O.Yt, spin, ytx, [v.copy() for v in x], g, a, 1, O.DP)
    lhs = sum(float(np.dot(p, q)) for p, q in zip(ya, x))
    mvals = np.concatenate([np.full(lmax + 1 - m, m) for m in range(lmax + 1)])
    c = np.where(mvals == 0, 1.0, 2.0)
    rhs = sum(float(np.sum(c * np.real(np.conj(p) * q))) for p, q in zip(alms, ytx))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


# --- geometry and mlim ------------------------------------------------------
def _geom_eq(g1, g2):
    """Same pairs in the same order; integers exact, doubles to 1e-13 (the
    reference is compiled with -ffast-math, which re-associates the theta
    formulas and the quadrature-weight sums at the 1e-15..1e-14 level)."""
    r1, n1 = O.geom_arrays(g1)
    r2, n2 = O.geom_arrays(g2)
    assert n1 == n2 and len(r1) == len(r2)
    for f in r1.dtype.names:
        a, b = r1[f], r2[f]
        if f.startswith("r2_") and f != "r2_nph":
            mask = r1["r2_nph"] > 0
            a, b = a[mask], b[mask]
        if a.dtype.kind == "f":
            np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-300, err_msg=f)
        else:
            np.testing.assert_array_equal(a, b, err_msg=f)


@pytest.mark.parametrize("maker", [
    lambda L: L.healpix(16), lambda L: L.healpix(5, np.linspace(0.5, 1.5, 10)),
    lambda L: L.iso("gauss", 33, 64), lambda L: L.iso("fejer1", 20, 41),
    lambda L: L.iso("cc", 21, 40), lambda L: L.iso("fejer2", 19, 38),
    lambda L: L.iso("mw", 17, 33)])
def test_geometry_matches_reference(oracle, ref, maker):
    _geom_eq(maker(oracle), maker(ref))


def test_mlim_matches_reference(oracle, ref):
    rng = np.random.default_rng(2)
    for _ in range(200):
        th = rng.uniform(0, math.pi)
        lmax = int(rng.integers(1, 5000))
        spin = int(rng.integers(0, 4))
        assert oracle.mlim(lmax, spin, math.sin(th), math.cos(th)) == \
            ref.mlim(lmax, spin, math.sin(th), math.cos(th))

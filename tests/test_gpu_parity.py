"""Parity of the HIP engine with the C restatement and the reference build.

Every case builds the same geometry and a_lm layout with each library's own
constructors, feeds identical seeded inputs through sharp_execute, and
compares all outputs: FP64 within 1e-12 relative L2 (the tolerance the
reference's own tests use for its golden values), FP32 within 1e-5.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from sht_cases import flip_theta, healpix, iso, rect, run_case, tri

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL64 = 1e-12
TOL32 = 1e-5


def compare(product, checker, *args, tol=TOL64, **kw):
    pa, pm = run_case(product, *args, **kw)
    ca, cm = run_case(checker, *args, **kw)
    errs = [rel_l2(x, y) for x, y in zip(pa + pm, ca + cm)]
    assert max(errs) < tol, errs
    return max(errs)


def test_sign_scale_goldens(product, gpu):
    """libsharp/sharp_testsuite.c:337-443 on the GPU, ntrans 1..3."""
    gold = json.load(open(os.path.join(HERE, "golden", "sign_scale.json")))
    lmax, nr, pp = gold["lmax"], gold["nrings"], gold["ppring"]
    npix = nr * pp
    nalm = (lmax + 1) * (lmax + 2) // 2
    for ntrans in (1, 2, 3):
        g = product.iso("gauss", nr, pp)
        flip_theta(g)
        a = product.triangular(lmax)
        for case in gold["cases"]:
            nmaps = len(case["maps"])
            na = 1 if (case["spin"] == 0 or case["type"] == O.DERIV1) else 2
            alms = [np.full(nalm, 1 + 1j) for _ in range(ntrans * na)]
            maps = [np.zeros(npix) for _ in range(ntrans * nmaps)]
            product.execute(case["type"], case["spin"], alms, maps, g, a, ntrans, O.DP)
            for t in range(ntrans):
                for k, want in enumerate(case["maps"]):
                    got = maps[t * nmaps + k][[0, npix // 2, npix - 1]]
                    np.testing.assert_allclose(got, want, rtol=max(case["rtol"]))
        product.destroy_geom(g)
        product.destroy_alm(a)


@pytest.mark.parametrize("spin", [0, 1, 2, 3])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_healpix_small(product, oracle, gpu, spin, jobtype):
    compare(product, oracle, healpix(8), tri(16), jobtype, spin, 1, O.DP)


@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM, O.Yt, O.WY])
@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("ntrans", [1, 3, 6])
def test_gauss_jobtypes_ntrans(product, oracle, gpu, jobtype, spin, ntrans):
    compare(product, oracle, iso("gauss", 25, 49), tri(24), jobtype, spin, ntrans, O.DP)


def test_deriv1(product, oracle, gpu):
    compare(product, oracle, healpix(8), tri(20), O.DERIV1, 1, 2, O.DP)


@pytest.mark.parametrize("extra", [O.ADD, O.REAL_HARMONICS, O.ADD | O.REAL_HARMONICS])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_flags(product, oracle, gpu, extra, jobtype):
    compare(product, oracle, healpix(8), tri(18), jobtype, 2, 1, O.DP | extra)
    compare(product, oracle, healpix(8), tri(18), jobtype, 0, 2, O.DP | extra)


@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_single_precision(product, oracle, gpu, jobtype):
    compare(product, oracle, healpix(8), tri(16), jobtype, 0, 1, 0, tol=TOL32)
    compare(product, oracle, healpix(8), tri(16), jobtype, 2, 1, 0, tol=TOL32)


@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_no_fft(product, oracle, gpu, jobtype):
    # one complex value per m per ring: nph = mmax+1
    lmax = 20
    compare(product, oracle, iso("gauss", 21, lmax + 1), tri(lmax), jobtype, 0, 1,
            O.DP | O.NO_FFT)
    compare(product, oracle, iso("gauss", 21, lmax + 1), tri(lmax), jobtype, 2, 1,
            O.DP | O.NO_FFT)


@pytest.mark.parametrize("kind", ["fejer1", "cc", "fejer2", "mw"])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_other_grids(product, oracle, gpu, kind, jobtype):
    compare(product, oracle, iso(kind, 33, 40), tri(16), jobtype, 0, 1, O.DP)
    compare(product, oracle, iso(kind, 33, 40), tri(16), jobtype, 2, 1, O.DP)


@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_rectangular_and_mmax(product, oracle, gpu, jobtype):
    compare(product, oracle, healpix(8), rect(20, 12), jobtype, 0, 1, O.DP)
    compare(product, oracle, healpix(8), tri(20, 7), jobtype, 2, 1, O.DP)


def _general_geom(lib):
    """Unequal rings: 1-pixel rings, an unpaired ring, strided and gapped
    pixel indices (the shapes pntsrc2alm and convolver_helper use)."""
    theta = np.array([0.3, np.pi - 0.3, 0.9, np.pi - 0.9, 1.4, 2.0, 1.2])
    nph = np.array([1, 1, 7, 7, 12, 5, 9])
    stride = np.array([1, 1, 2, 2, 1, 3, 1])
    ofs = np.zeros(len(nph), np.int64)
    pos = 3
    for i in range(len(nph)):
        ofs[i] = pos
        pos += nph[i] * stride[i] + 2
    phi0 = np.array([0.1, 0.0, 0.5, 0.25, 0.0, 1.0, 0.3])
    wgt = np.linspace(0.5, 1.5, len(nph))
    return lib.general_geom(nph, ofs, stride, phi0, theta, wgt)


@pytest.mark.parametrize("spin", [0, 1, 2])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM, O.Yt, O.WY])
def test_general_geometry(product, oracle, gpu, spin, jobtype):
    compare(product, oracle, _general_geom, tri(10), jobtype, spin, 2, O.DP)


def _fft_class_geom(lib):
    """Rings whose FFTs take the ring stage's plans up to length 8192:
    Bluestein with M = 8192 (passes 8, 16, 8, 8; even nph 8156, 6010, 4100 and
    odd nph 4093), M = 4096 (nph 2062) and M = 2048 (nph 1990), radix 8 only
    (nph 8192) and mixed radices (nph 6000), with nph above 2 mmax + 1 (direct
    fold) and below it (staged fold)."""
    nph = np.array([8156, 4093, 6010, 4100, 8192, 6000, 2062, 1990])
    theta = np.linspace(0.2, 2.9, len(nph))
    stride = np.ones(len(nph), np.int64)
    ofs = np.concatenate([[0], np.cumsum(nph)[:-1]]).astype(np.int64)
    phi0 = np.linspace(0.0, 0.7, len(nph))
    wgt = np.linspace(0.5, 1.5, len(nph))
    return lib.general_geom(nph, ofs, stride, phi0, theta, wgt)


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_ring_fft_classes(product, oracle, gpu, spin, jobtype):
    compare(product, oracle, _fft_class_geom, tri(2100), jobtype, spin, 1, O.DP)


def _fft_alias_geom(lib):
    """Rings of the classes of 2048 entries and more whose phases, at mmax
    2100, fold with at most two aliases per bin (n <= mmax < 2n): even nph
    2062 (Bluestein M = 4096), 1990 (Bluestein M = 2048), 2100 = mmax (mixed
    radices; bin 0 takes m = 0 and m = n) and 2058 (N = 1029, radices 3 and
    7), odd nph 2025 (radices 3, 5) and 2075 (Bluestein M = 8192)."""
    nph = np.array([2062, 1990, 2100, 2058, 2025, 2075])
    theta = np.linspace(0.3, 2.8, len(nph))
    stride = np.ones(len(nph), np.int64)
    ofs = np.concatenate([[0], np.cumsum(nph)[:-1]]).astype(np.int64)
    phi0 = np.array([0.2, 0.0, 0.45, 0.1, 0.0, 0.6])
    wgt = np.linspace(0.7, 1.3, len(nph))
    return lib.general_geom(nph, ofs, stride, phi0, theta, wgt)


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("dp", [True, False])
def test_ring_fft_two_alias_fold(product, oracle, gpu, spin, dp):
    compare(product, oracle, _fft_alias_geom, tri(2100), O.ALM2MAP, spin, 2,
            O.DP if dp else 0, tol=TOL64 if dp else TOL32)


def _long_ring_geom(lib):
    """Rings longer than the largest LDS FFT class (8192 complex values):
    nph 16388 (N = 8194 = 2*17*241, Bluestein length 16384), 32768 (N =
    16384) and the odd 16385; they take the direct-sum ring kernels."""
    nph = np.array([16388, 32768, 16385, 16388])
    theta = np.array([0.4, 1.1, 1.9, np.pi - 0.4])
    stride = np.ones(len(nph), np.int64)
    ofs = np.concatenate([[0], np.cumsum(nph)[:-1]]).astype(np.int64)
    phi0 = np.array([0.0, 0.3, 1.0, 0.0])
    wgt = np.linspace(0.5, 1.5, len(nph))
    return lib.general_geom(nph, ofs, stride, phi0, theta, wgt)


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_rings_beyond_fft_classes(product, oracle, gpu, spin, jobtype):
    # every ring longer than 2 mmax + 1: the phases are padded, not folded
    compare(product, oracle, _long_ring_geom, tri(2100), jobtype, spin, 1, O.DP)


@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_rings_beyond_fft_classes_folded(product, oracle, gpu, jobtype):
    # mmax 8300: the rings of 16388 and 16385 pixels are shorter than
    # 2 mmax + 1, so their phases fold onto the ring's bins
    compare(product, oracle, _long_ring_geom, tri(8300), jobtype, 0, 1, O.DP)


def test_packed(product, oracle, gpu):
    def packed(lib):
        lmax = 12
        mval = np.arange(lmax + 1)
        # m=0 real-only block of lmax+1 reals, then complex blocks
        mvs = np.zeros(lmax + 1, np.int64)
        pos = lmax + 1
        for m in range(1, lmax + 1):
            mvs[m] = pos - 2 * m      # real-unit start so that l=m lands at pos
            pos += 2 * (lmax - m + 1)
        return lib.general_alm(lmax, mval, mvs, 1, O.PACKED)
    for jt in (O.ALM2MAP, O.MAP2ALM):
        compare(product, oracle, healpix(4), packed, jt, 0, 1, O.DP)
        compare(product, oracle, healpix(4), packed, jt, 2, 1, O.DP)


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_healpix_vs_reference(product, ref, gpu, spin, jobtype):
    compare(product, ref, healpix(32), tri(64), jobtype, spin, 1, O.DP)


@pytest.mark.parametrize("spin", [0, 30])
def test_high_spin_vs_reference(product, ref, gpu, spin):
    for jt in (O.ALM2MAP, O.MAP2ALM):
        compare(product, ref, iso("gauss", 64, 127), tri(63), jt, spin, 1, O.DP)


def test_cfg1_spin0_vs_reference(product, ref, gpu):
    """HEALPix nside 256, lmax 512 spin 0 in both directions."""
    for jt in (O.ALM2MAP, O.MAP2ALM):
        compare(product, ref, healpix(256), tri(512), jt, 0, 1, O.DP)


@pytest.mark.parametrize("spin", [0, 2, 3])
def test_fp64_analysis_reduction_kernels(product, oracle, gpu, spin, monkeypatch):
    """SHARP_HIP_FP64_ANALYSIS=reduce selects the FP64 analysis kernels that sum
    over ring pairs with lane reductions instead of matrix cores; both
    directions of analysis (map2alm, Y^T) match the oracle to 1e-12."""
    monkeypatch.setenv("SHARP_HIP_FP64_ANALYSIS", "reduce")
    for jt in (O.MAP2ALM, O.Yt):
        compare(product, oracle, healpix(16), tri(32), jt, spin, 3, O.DP)
        compare(product, oracle, iso("gauss", 25, 49), tri(24), jt, spin, 5, O.DP)

"""Device-resident Jacobi analysis (map2alm_iter, Healpix_cxx/
alm_healpix_tools.cc:73-85, 239-265) and transforms with per-l factors on the
a_lm side (smoothWithGauss + alm2map, Healpix_cxx/alm_powspec_tools.cc:250-283,
Healpix_cxx/smoothing_cxx_module.cc:79-81).

The checker runs the same iteration with the oracle's own sharp_execute
(map2alm, alm2map, subtract, map2alm with SHARP_ADD) and scales a_lm in numpy;
FP64 results agree to 1e-12 relative L2, the mixed form to 1e-6.  A property
independent of the oracle: on a band-limited HEALPix map the iteration error
|a_iter - a_true| falls with every iteration.
"""
import numpy as np
import pytest
import torch

import oracle as O
import sharp_hip as S
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOLMIX = 1e-6


def _lm_of(ainfo_c):
    """l of every complex entry of a triangular a_lm array (numpy)."""
    ai = ainfo_c.contents
    ls = np.zeros(max(ai.mvstart[mi] + ai.lmax * ai.stride + 1 for mi in range(ai.nm)),
                  dtype=np.int64)
    for mi in range(ai.nm):
        m = ai.mval[mi]
        idx = ai.mvstart[mi] + np.arange(m, ai.lmax + 1) * ai.stride
        ls[idx] = np.arange(m, ai.lmax + 1)
    return ls


def _random_alms(rng, nalm, n, lmax, ls, mzero):
    out = []
    for _ in range(n):
        a = rng.standard_normal(nalm) + 1j * rng.standard_normal(nalm)
        a[mzero] = a[mzero].real
        out.append(a / (1.0 + ls))
    return out


def _oracle_iter(orc, spin, maps, g, a, ntrans, niter, nalm):
    nc = 1 if spin == 0 else 2
    alms = [np.zeros(nalm, np.complex128) for _ in range(ntrans * nc)]
    orc.execute(O.MAP2ALM, spin, alms, [m.copy() for m in maps], g, a, ntrans, O.DP)
    for _ in range(niter):
        y = [np.zeros_like(m) for m in maps]
        orc.execute(O.ALM2MAP, spin, alms, y, g, a, ntrans, O.DP)
        r = [m - yy for m, yy in zip(maps, y)]
        orc.execute(O.MAP2ALM, spin, alms, r, g, a, ntrans, O.DP | O.ADD)
    return alms


def _to_dev(xs, dev):
    return [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in xs]


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("niter", [0, 1, 3])
def test_map2alm_iter_matches_oracle(oracle, gpu, spin, niter, mixed):
    nside, lmax, ntrans = 16, 40, 2
    nc = 1 if spin == 0 else 2
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    g, a = oracle.healpix(nside), oracle.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    rng = np.random.default_rng(7 + spin + 10 * niter)
    maps = [rng.uniform(-1, 1, geom.npix) for _ in range(ntrans * nc)]
    want = _oracle_iter(oracle, spin, maps, g, a, ntrans, niter, nalm)
    dmaps = _to_dev(maps, gpu)
    dalms = [torch.zeros(nalm, dtype=torch.complex128, device=gpu) for _ in want]
    flags = S.DP | (S.MIXED_FP32 if mixed else 0)
    S.map2alm_iter_device(spin, dalms, dmaps, geom, ainfo, ntrans, flags, niter)
    torch.cuda.synchronize()
    errs = [rel_l2(x.cpu().numpy(), y) for x, y in zip(dalms, want)]
    assert max(errs) < (TOLMIX if mixed else TOL64), errs
    if mixed:
        assert min(errs) > 1e-12, errs
    # the input maps are left as they were
    assert all(torch.equal(x.cpu(), torch.from_numpy(y)) for x, y in zip(dmaps, maps))
    oracle.destroy_geom(g)
    oracle.destroy_alm(a)


def test_map2alm_iter_gauss_general_and_float(oracle, gpu):
    """A Gauss-Legendre grid (the iteration is exact after step 0 there) and
    single-precision maps."""
    lmax, ntrans = 23, 3
    geom, ainfo = S.Geometry.iso("gauss", lmax + 1, 2 * lmax + 1), S.AlmInfo.triangular(lmax)
    g, a = oracle.iso("gauss", lmax + 1, 2 * lmax + 1), oracle.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    rng = np.random.default_rng(11)
    maps = [rng.uniform(-1, 1, geom.npix) for _ in range(ntrans * 2)]
    want = _oracle_iter(oracle, 2, maps, g, a, ntrans, 2, nalm)
    dalms = [torch.zeros(nalm, dtype=torch.complex128, device=gpu) for _ in want]
    S.map2alm_iter_device(2, dalms, _to_dev(maps, gpu), geom, ainfo, ntrans, S.DP, 2)
    torch.cuda.synchronize()
    errs = [rel_l2(x.cpu().numpy(), y) for x, y in zip(dalms, want)]
    assert max(errs) < TOL64, errs
    # float maps and a_lm (no SHARP_DP): the oracle in float as well
    maps32 = [m.astype(np.float32) for m in maps[:2]]
    want32 = [np.zeros(nalm, np.complex64) for _ in range(2)]
    orc_in = [m.copy() for m in maps32]
    oracle.execute(O.MAP2ALM, 2, want32, orc_in, g, a, 1, 0)
    y = [np.zeros_like(m) for m in maps32]
    oracle.execute(O.ALM2MAP, 2, want32, y, g, a, 1, 0)
    r = [m - yy for m, yy in zip(maps32, y)]
    oracle.execute(O.MAP2ALM, 2, want32, r, g, a, 1, O.ADD)
    d32 = [torch.zeros(nalm, dtype=torch.complex64, device=gpu) for _ in range(2)]
    S.map2alm_iter_device(2, d32, _to_dev(maps32, gpu), geom, ainfo, 1, 0, 1)
    torch.cuda.synchronize()
    errs = [rel_l2(x.cpu().numpy(), y) for x, y in zip(d32, want32)]
    assert max(errs) < 1e-5, errs
    oracle.destroy_geom(g)
    oracle.destroy_alm(a)


def test_iteration_converges_on_band_limited_map(gpu):
    """A map synthesised from a_lm with lmax <= 1.5 Nside: each Jacobi step
    brings map2alm closer to the true a_lm (HEALPix quadrature is not exact,
    the iteration removes its error)."""
    nside, lmax = 32, 48
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    ls = _lm_of(ainfo.ptr)
    mzero = np.arange(lmax + 1)
    rng = np.random.default_rng(5)
    atrue = _to_dev(_random_alms(rng, nalm, 1, lmax, ls, mzero), gpu)
    m = [torch.zeros(geom.npix, dtype=torch.float64, device=gpu)]
    S.execute_device(S.ALM2MAP, 0, atrue, m, geom, ainfo, 1, S.DP)
    errs = []
    for niter in range(4):
        out = [torch.zeros(nalm, dtype=torch.complex128, device=gpu)]
        S.map2alm_iter_device(0, out, m, geom, ainfo, 1, S.DP, niter)
        torch.cuda.synchronize()
        errs.append(rel_l2(out[0].cpu().numpy(), atrue[0].cpu().numpy()))
    assert all(e2 < 0.5 * e1 for e1, e2 in zip(errs, errs[1:])), errs


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("spin", [0, 2])
def test_scaled_synthesis_and_analysis(oracle, gpu, spin, mixed):
    """Y (f_l a) and f_l (Y^T W m), with SHARP_ADD for the analysis; the
    caller's a_lm are not modified by the synthesis."""
    nside, lmax, ntrans = 16, 36, 2
    nc = 1 if spin == 0 else 2
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    g, a = oracle.healpix(nside), oracle.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    ls = _lm_of(ainfo.ptr)
    rng = np.random.default_rng(3 + spin)
    fl = S.gauss_beam(lmax, np.radians(90.0 / 60.0), pol=spin != 0)
    alms = _random_alms(rng, nalm, ntrans * nc, lmax, ls, np.arange(lmax + 1))
    flags = S.DP | (S.MIXED_FP32 if mixed else 0)
    tol = TOLMIX if mixed else TOL64
    # synthesis
    want = [np.zeros(geom.npix) for _ in range(ntrans * nc)]
    oracle.execute(O.ALM2MAP, spin, [x * fl[ls] for x in alms], want, g, a, ntrans, O.DP)
    dal = _to_dev(alms, gpu)
    dmp = [torch.zeros(geom.npix, dtype=torch.float64, device=gpu) for _ in want]
    S.execute_scaled_device(S.ALM2MAP, spin, dal, dmp, geom, ainfo, ntrans, flags, fl)
    torch.cuda.synchronize()
    errs = [rel_l2(x.cpu().numpy(), y) for x, y in zip(dmp, want)]
    assert max(errs) < tol, errs
    assert all(torch.equal(x.cpu(), torch.from_numpy(y)) for x, y in zip(dal, alms))
    # analysis, accumulated onto existing a_lm
    maps = [rng.uniform(-1, 1, geom.npix) for _ in range(ntrans * nc)]
    base = _random_alms(rng, nalm, ntrans * nc, lmax, ls, np.arange(lmax + 1))
    got = [np.zeros(nalm, np.complex128) for _ in base]
    oracle.execute(O.MAP2ALM, spin, got, [m.copy() for m in maps], g, a, ntrans, O.DP)
    want = [b + fl[ls] * x for b, x in zip(base, got)]
    dal = _to_dev(base, gpu)
    S.execute_scaled_device(S.MAP2ALM, spin, dal, _to_dev(maps, gpu), geom, ainfo, ntrans,
                            flags | S.ADD, fl)
    torch.cuda.synchronize()
    errs = [rel_l2(x.cpu().numpy(), y) for x, y in zip(dal, want)]
    assert max(errs) < tol, errs
    oracle.destroy_geom(g)
    oracle.destroy_alm(a)


def test_smoothing_matches_composition(oracle, gpu):
    """smoothing_cxx's core for T and Q/U: map2alm_iter, smoothWithGauss,
    alm2map, against the same steps on the oracle."""
    nside, lmax, niter = 16, 32, 2
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    g, a = oracle.healpix(nside), oracle.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    ls = _lm_of(ainfo.ptr)
    fwhm = np.radians(120.0 / 60.0)
    rng = np.random.default_rng(21)
    for spin in (0, 2):
        nc = 1 if spin == 0 else 2
        maps = [rng.uniform(-1, 1, geom.npix) for _ in range(nc)]
        al = _oracle_iter(oracle, spin, maps, g, a, 1, niter, nalm)
        fl = S.gauss_beam(lmax, fwhm, pol=spin != 0)
        want = [np.zeros(geom.npix) for _ in range(nc)]
        oracle.execute(O.ALM2MAP, spin, [x * fl[ls] for x in al], want, g, a, 1, O.DP)
        dm = _to_dev(maps, gpu)
        dal = [torch.zeros(nalm, dtype=torch.complex128, device=gpu) for _ in range(nc)]
        S.smoothing_device(dm, geom, ainfo, spin, fwhm, niter, dal)
        torch.cuda.synchronize()
        errs = [rel_l2(x.cpu().numpy(), y) for x, y in zip(dm, want)]
        assert max(errs) < TOL64, (spin, errs)
    oracle.destroy_geom(g)
    oracle.destroy_alm(a)

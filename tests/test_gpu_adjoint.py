"""Differentiation form on the GPU: the adjoint entry points and the
torch.autograd harness over them.

* adjoint identity <Y a, x> = sum_m c_m Re<a_m, (Y^T x)_m> (c_0=1, c_m>0=2)
  for the FP64 and the mixed form, spins 0 and 2;
* the chi^2 workload: HEALPix Nside 512, lmax 1024, TQU; d = alm2map(a_true)
  + N(0, 0.1^2), w ~ U[0.5, 2] and w = 0 on the |b| < 10 deg band;
  chi^2 = sum_pix w (alm2map(a) - d)^2.  Its autograd gradient is checked by
  central differences (h = 1e-6) on 64 random real a_lm entries (tolerance
  1e-6 relative plus the rounding of the loss sums over 2h: the loss is
  quadratic, so the differences are exact up to rounding) and against the reference build's own adjoint (SHARP_Yt) at
  1e-12 relative L2.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
import sharp_hip as S
import sht_autograd as A

pytestmark = pytest.mark.gpu


def _ring_theta(geom):
    """theta of every pixel of a ring geometry (host numpy)."""
    gi = geom.ptr.contents
    th = np.zeros(geom.npix)
    for i in range(gi.npairs):
        p = gi.pair[i]
        for r in (p.r1, p.r2):
            if r.nph > 0:
                th[r.ofs + np.arange(r.nph) * r.stride] = r.theta
    return th


def _cm(ainfo, nalm):
    return A.m_weights(ainfo, nalm, "cpu").numpy()


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("spin", [0, 2])
def test_adjoint_identity(gpu, spin, mixed):
    nside, lmax = 64, 128
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    nc = 1 if spin == 0 else 2
    flags = S.DP | (S.MIXED_FP32 if mixed else 0)
    g = torch.Generator(device=gpu).manual_seed(3)
    a = [torch.complex(torch.randn(nalm, generator=g, device=gpu, dtype=torch.float64),
                       torch.randn(nalm, generator=g, device=gpu, dtype=torch.float64))
         for _ in range(nc)]
    x = [torch.randn(geom.npix, generator=g, device=gpu, dtype=torch.float64) for _ in range(nc)]
    ya = [torch.zeros(geom.npix, dtype=torch.float64, device=gpu) for _ in range(nc)]
    ytx = [torch.zeros(nalm, dtype=torch.complex128, device=gpu) for _ in range(nc)]
    S.execute_device(S.ALM2MAP, spin, a, ya, geom, ainfo, 1, flags)
    S.alm2map_adjoint_device(spin, ytx, x, geom, ainfo, 1, flags)
    torch.cuda.synchronize()
    cm = torch.from_numpy(_cm(ainfo, nalm)).to(gpu)
    lhs = sum(float((yi * xi).sum()) for yi, xi in zip(ya, x))
    # m = 0: only Re a_l0 enters the map
    rhs = 0.0
    for ai, ti in zip(a, ytx):
        re = ai.real * ti.real
        im = torch.where(cm == 1.0, torch.zeros_like(re), ai.imag * ti.imag)
        rhs += float((cm * (re + im)).sum())
    if mixed:
        # the per-entry error (~1e-7 relative) of the FP32 contraction enters
        # the inner product with the scale ||Y a|| ||x|| (Cauchy-Schwarz)
        scale = math.sqrt(sum(float((yi * yi).sum()) for yi in ya) *
                          sum(float((xi * xi).sum()) for xi in x))
        assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs, scale)
    else:
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_chi2_gradient_nside512_tqu(gpu, ref):
    nside, lmax = 512, 1024
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    npix, nalm = geom.npix, S.nalm_triangular(lmax)
    s0, s2 = A.SHT(geom, ainfo, 0), A.SHT(geom, ainfo, 2)
    rng = np.random.default_rng(2024)
    l_of = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    cl = 1.0 / (l_of * (l_of + 1.0) + 1.0)

    def draw():
        a = np.stack([rng.normal(size=nalm), rng.normal(size=nalm)], axis=1)
        a *= np.sqrt(cl / 2)[:, None]
        return torch.tensor(a, device=gpu)
    a_true = [draw(), draw(), draw()]
    with torch.no_grad():
        mt = s0.alm2map(a_true[0])[0]
        mq, mu = s2.alm2map(a_true[1], a_true[2])
    d = [m + 0.1 * torch.tensor(rng.normal(size=npix), device=gpu) for m in (mt, mq, mu)]
    theta = _ring_theta(geom)
    band = np.abs(0.5 * math.pi - theta) < math.radians(10.0)
    w = [torch.tensor(np.where(band, 0.0, rng.uniform(0.5, 2.0, size=npix)), device=gpu)
         for _ in range(3)]
    a = [x + 0.01 * draw() for x in a_true]
    for x in a:
        x.requires_grad_(True)

    def chi2(t, e, b):
        m0 = s0.alm2map(t)[0]
        m1, m2 = s2.alm2map(e, b)
        return sum((wi * (mi - di) ** 2).sum() for wi, mi, di in zip(w, (m0, m1, m2), d))

    L = chi2(*a)
    L.backward()
    grads = [x.grad.detach().clone() for x in a]

    # central differences on 64 random real entries
    h = 1e-6
    for _ in range(64):
        c = int(rng.integers(3))
        i = int(rng.integers(nalm))
        p = int(rng.integers(2))
        if p == 1 and i <= lmax:   # m = 0 (the first lmax+1 entries): Im is not used
            p = 0
        with torch.no_grad():
            ap = [x.detach().clone() for x in a]
            am = [x.detach().clone() for x in a]
            ap[c][i, p] += h
            am[c][i, p] -= h
            fd = (chi2(*ap).item() - chi2(*am).item()) / (2 * h)
        g = grads[c][i, p].item()
        # rounding of the two loss sums, divided by 2h, bounds the difference
        # quotient's own error
        tol = 1e-6 * abs(g) + 50 * L.item() * 2.2e-16 / h
        assert abs(fd - g) <= tol, (c, i, p, fd, g, tol)

    # against the reference build's adjoint: grad = c_m * Yt(2 w r)
    with torch.no_grad():
        m0 = s0.alm2map(a[0])[0]
        m1, m2 = s2.alm2map(a[1], a[2])
        r = [(2 * wi * (mi - di)).cpu().numpy() for wi, mi, di in zip(w, (m0, m1, m2), d)]
    rg, ra = ref.healpix(nside), ref.triangular(lmax)
    cm = _cm(ainfo, nalm)
    out_t = [np.zeros(nalm, np.complex128)]
    out_eb = [np.zeros(nalm, np.complex128), np.zeros(nalm, np.complex128)]
    ref.execute(O.Yt, 0, out_t, [r[0]], rg, ra, 1, O.DP)
    ref.execute(O.Yt, 2, out_eb, [r[1], r[2]], rg, ra, 1, O.DP)
    for k, o in enumerate(out_t + out_eb):
        want = np.stack([o.real * cm, np.where(cm == 1.0, 0.0, o.imag * cm)], axis=1)
        got = grads[k].cpu().numpy()
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < 1e-12, (k, err)

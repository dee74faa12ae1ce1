"""The autograd harness's gradient convention, checked on the CPU: the
transforms come from the oracle (as the engine object the harness accepts),
so this exercises the c_m factors, the m=0 handling and the autograd
plumbing without a GPU; the GPU tests repeat the checks on the HIP engine.

Checks: central finite differences (h=1e-6) of a weighted chi^2 through
alm2map for every real a_lm degree of freedom, and of an a_lm-domain loss
through map2alm for map pixels.  Tolerance 1e-7 relative (the losses are
quadratic, so central differences are exact up to rounding)."""
import numpy as np
import pytest
import torch

import oracle as O
import sht_autograd as A


class _Geom:
    def __init__(self, lib, nside):
        self.lib, self.ptr = lib, lib.healpix(nside)
        self._npix = 12 * nside * nside

    @property
    def npix(self):
        return self._npix


class _Alm:
    def __init__(self, lib, lmax):
        self.lib, self.ptr = lib, lib.triangular(lmax)


class OracleEngine:
    """alm2map / Y^T / map2alm / W Y on CPU tensors through the oracle."""

    def __init__(self, lib, geom, ainfo, spin):
        self.lib, self.g, self.a, self.spin = lib, geom.ptr, ainfo.ptr, spin

    def _run(self, jt, alms, maps):
        an = [a.detach().numpy().copy() for a in alms]
        mn = [m.detach().numpy().copy() for m in maps]
        self.lib.execute(jt, self.spin, an, mn, self.g, self.a, 1, O.DP)
        return an, mn

    def synth(self, alms, maps):
        _, mn = self._run(O.ALM2MAP, alms, maps)
        for m, v in zip(maps, mn):
            m.copy_(torch.from_numpy(v))

    def synth_adjoint(self, maps, alms):
        an, _ = self._run(O.Yt, alms, maps)
        for a, v in zip(alms, an):
            a.copy_(torch.from_numpy(v))

    def anal(self, maps, alms):
        an, _ = self._run(O.MAP2ALM, alms, maps)
        for a, v in zip(alms, an):
            a.copy_(torch.from_numpy(v))

    def anal_adjoint(self, alms, maps):
        _, mn = self._run(O.WY, alms, maps)
        for m, v in zip(maps, mn):
            m.copy_(torch.from_numpy(v))


def _sht(oracle, spin, nside=4, lmax=8):
    g, a = _Geom(oracle, nside), _Alm(oracle, lmax)
    return A.SHT(g, a, spin, engine=OracleEngine(oracle, g, a, spin), device="cpu")


@pytest.mark.parametrize("spin", [0, 2])
def test_alm2map_gradient_fd(oracle, spin):
    sht = _sht(oracle, spin)
    rng = np.random.default_rng(7 + spin)
    nc = 1 if spin == 0 else 2
    alms = [torch.tensor(rng.normal(size=(sht.nalm, 2)), requires_grad=True)
            for _ in range(nc)]
    d = [torch.tensor(rng.normal(size=sht.npix)) for _ in range(nc)]
    w = [torch.tensor(rng.uniform(0.5, 2.0, size=sht.npix)) for _ in range(nc)]

    def loss(*a):
        maps = sht.alm2map(*a)
        return sum((wi * (mi - di) ** 2).sum() for wi, mi, di in zip(w, maps, d))

    L = loss(*alms)
    L.backward()
    h = 1e-6
    for c in range(nc):
        g = alms[c].grad.numpy()
        for i in range(sht.nalm):
            for p in range(2):
                ap = [x.detach().clone() for x in alms]
                am = [x.detach().clone() for x in alms]
                ap[c][i, p] += h
                am[c][i, p] -= h
                fd = (loss(*ap).item() - loss(*am).item()) / (2 * h)
                assert abs(fd - g[i, p]) <= 1e-7 * max(1.0, abs(fd)) + 1e-6, (c, i, p, fd, g[i, p])


@pytest.mark.parametrize("spin", [0, 2])
def test_map2alm_gradient_fd(oracle, spin):
    sht = _sht(oracle, spin, nside=2, lmax=4)
    rng = np.random.default_rng(11 + spin)
    nc = 1 if spin == 0 else 2
    maps = [torch.tensor(rng.normal(size=sht.npix), requires_grad=True) for _ in range(nc)]
    t = [torch.tensor(rng.normal(size=(sht.nalm, 2))) for _ in range(nc)]

    def loss(*m):
        alms = sht.map2alm(*m)
        return sum(((a - ti) ** 2).sum() for a, ti in zip(alms, t))

    L = loss(*maps)
    L.backward()
    h = 1e-6
    for c in range(nc):
        g = maps[c].grad.numpy()
        for i in range(sht.npix):
            mp = [x.detach().clone() for x in maps]
            mm = [x.detach().clone() for x in maps]
            mp[c][i] += h
            mm[c][i] -= h
            fd = (loss(*mp).item() - loss(*mm).item()) / (2 * h)
            assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)) + 1e-6, (c, i, fd, g[i])

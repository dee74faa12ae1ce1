"""Seeded synthetic inputs of the benchmark and parity workloads (torch, on
the GPU): Gaussian a_lm in the Healpix triangular layout drawn from CMB-like
spectra, with the correlated T/E draw of Healpix_cxx/alm_powspec_tools.cc:
71-133 (m = 0 real with variance C_l; m > 0 with Re and Im of variance
C_l / 2)."""
import math

import numpy as np


def spectra(lmax, fwhm_arcmin=5.0):
    """C_l^TT = 1/(l(l+1)+1) times the window b_l^2 of a Gaussian beam of the
    given FWHM (none for 0), C^EE = 0.1 C^TT, C^BB = 0.01 C^TT, TE correlation
    coefficient 0.5."""
    l = np.arange(lmax + 1, dtype=np.float64)
    sigma = fwhm_arcmin / 60.0 * math.pi / 180.0 / math.sqrt(8.0 * math.log(2.0))
    bl2 = np.exp(-l * (l + 1) * sigma * sigma)
    tt = bl2 / (l * (l + 1) + 1)
    ee = 0.1 * tt
    bb = 0.01 * tt
    te = 0.5 * np.sqrt(tt * ee)
    return tt, ee, bb, te


def make_alms(torch, lmax, batch, seed, dev, fwhm_arcmin=5.0):
    """`batch` realisations of (T, E, B) a_lm: three lists of complex128
    tensors on `dev` in the triangular layout of lmax = mmax."""
    tt, ee, bb, te = spectra(lmax, fwhm_arcmin)
    m_of = np.concatenate([np.full(lmax + 1 - m, m) for m in range(lmax + 1)])
    l_of = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    l_t = torch.from_numpy(l_of).to(dev)
    m0 = torch.from_numpy(m_of == 0).to(dev)
    ctt = torch.from_numpy(tt).to(dev)[l_t]
    cee = torch.from_numpy(ee).to(dev)[l_t]
    cbb = torch.from_numpy(bb).to(dev)[l_t]
    cte = torch.from_numpy(te).to(dev)[l_t]
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    n = len(l_of)

    def cgauss():
        re = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        im = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        s = torch.where(m0, torch.ones_like(re), torch.full_like(re, math.sqrt(0.5)))
        im = torch.where(m0, torch.zeros_like(im), im)
        return torch.complex(re * s, im * s)

    T, E, B = [], [], []
    for _ in range(batch):
        g1, g2, g3 = cgauss(), cgauss(), cgauss()
        at = torch.sqrt(ctt) * g1
        ae = (cte / torch.sqrt(ctt)) * g1 + torch.sqrt(torch.clamp(cee - cte * cte / ctt, min=0)) * g2
        ab = torch.sqrt(cbb) * g3
        T.append(at.contiguous())
        E.append(ae.contiguous())
        B.append(ab.contiguous())
    return T, E, B


def spin0_alms(torch, lmax, count, gen, dev):
    """`count` spin-0 a_lm with C_l = 1/(l(l+1)+1) drawn from generator
    `gen` (m = 0 real)."""
    l_of = torch.cat([torch.arange(m, lmax + 1, device=dev) for m in range(lmax + 1)]).double()
    sd = torch.sqrt(1.0 / (l_of * (l_of + 1) + 1) / 2)
    out = []
    for _ in range(count):
        re = torch.randn(l_of.shape, generator=gen, device=dev, dtype=torch.float64)
        im = torch.randn(l_of.shape, generator=gen, device=dev, dtype=torch.float64)
        im[: lmax + 1] = 0.0
        out.append(torch.complex(re * sd, im * sd))
    return out

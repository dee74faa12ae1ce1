"""Case runner shared by the parity tests: builds the same geometry and a_lm
layout with each library's own constructors, feeds identical inputs, and
returns the outputs."""
import math

import numpy as np

import oracle as O


def geom_npix(g):
    """1 + largest pixel index used by the geometry."""
    gi = g.contents
    hi = 0
    for i in range(gi.npairs):
        p = gi.pair[i]
        for r in (p.r1, p.r2):
            if r.nph > 0:
                hi = max(hi, r.ofs + (r.nph - 1) * r.stride + 1, r.ofs + 1)
    return hi


def ncomp(jobtype, spin):
    if jobtype == O.DERIV1:
        return 1, 2
    return (1, 1) if spin == 0 else (2, 2)


def flip_theta(g):
    """Emulate the old Gaussian-grid orientation (libsharp/sharp_testsuite.c:347-355)."""
    gi = g.contents
    for i in range(gi.npairs):
        p = gi.pair[i]
        p.r1.cth = -p.r1.cth
        p.r2.cth = -p.r2.cth
        p.r1.theta = math.pi - p.r1.theta
        p.r2.theta = math.pi - p.r2.theta


def alm_len(a):
    ai = a.contents
    hi = 0
    packed = ai.flags & O.PACKED
    for mi in range(ai.nm):
        m = ai.mval[mi]
        for l in (m, ai.lmax):
            if packed:
                idx = ai.mvstart[mi] + l * ai.stride * (1 if m == 0 else 2)
                hi = max(hi, idx + (1 if m == 0 else 2))
            else:
                hi = max(hi, 2 * (ai.mvstart[mi] + l * ai.stride) + 2)
    return hi  # in reals


def make_inputs(jobtype, spin, ntrans, nalm_reals, npix, nofft, dp, rng, with_out=False):
    """Random inputs; returns (alms, maps) as lists of numpy arrays."""
    na, nm = ncomp(jobtype, spin)
    fdt = np.float64 if dp else np.float32
    alms = [rng.uniform(-1, 1, nalm_reals).astype(fdt) for _ in range(ntrans * na)]
    mlen = 2 * npix if nofft else npix
    maps = [rng.uniform(-1, 1, mlen).astype(fdt) for _ in range(ntrans * nm)]
    return alms, maps


def run_case(lib, geom_fn, alm_fn, jobtype, spin, ntrans, flags, seed=0,
             add=False, geom_post=None):
    """Runs one sharp_execute; returns (alms, maps) after the call."""
    g = geom_fn(lib)
    if geom_post is not None:
        geom_post(g)
    a = alm_fn(lib)
    dp = bool(flags & O.DP)
    nofft = bool(flags & O.NO_FFT)
    rng = np.random.default_rng(seed)
    alms, maps = make_inputs(jobtype, spin, ntrans, alm_len(a), geom_npix(g), nofft,
                             dp, rng)
    synth = jobtype in (O.ALM2MAP, O.WY, O.DERIV1)
    if synth and not add:
        # outputs start from garbage the call must overwrite (ring pixels only)
        pass
    lib.execute(jobtype, spin, alms, maps, g, a, ntrans, flags | (O.ADD if add else 0))
    lib.destroy_geom(g)
    lib.destroy_alm(a)
    return alms, maps


def healpix(nside):
    return lambda lib: lib.healpix(nside)


def iso(kind, nrings, nphi):
    return lambda lib: lib.iso(kind, nrings, nphi)


def tri(lmax, mmax=None):
    return lambda lib: lib.triangular(lmax, mmax)


def rect(lmax, mmax=None):
    return lambda lib: lib.rectangular(lmax, mmax)

"""The allocation-free entry point sharp_hip_execute_device_noalloc
(include/sharp_hip.h; no counterpart in the reference): through the C ABI,
its outputs equal those of sharp_hip_execute_device on the same inputs, for
two calls in a row on one GPU, the second with a smaller job (the scratch
area keeps the first call's contents).  Outputs start from random values, so
any entry a call fails to write shows up."""
import numpy as np
import pytest
import torch

import sharp_hip as S
from conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [  # (job type, spin, form flags)
    (S.ALM2MAP, 0, 0), (S.MAP2ALM, 0, 0), (S.ALM2MAP, 2, 0), (S.MAP2ALM, 2, 0),
    (S.ALM2MAP_DERIV1, 1, 0), (S.ALM2MAP, 2, S.MIXED_BF16X3), (S.MAP2ALM, 2, S.MIXED_BF16X3),
    (S.ALM2MAP, 0, S.MIXED_FP32), (S.MAP2ALM, 0, S.MIXED_FP32), (S.Yt, 2, 0), (S.WY, 0, 0),
]
SIZES = [(32, 64, 3), (16, 32, 1)]   # (nside, lmax, ntrans): larger first


def _inputs(dev, rng, geom, ainfo, jt, spin, nt):
    na = 1 if (spin == 0 or jt == S.ALM2MAP_DERIV1) else 2
    nm = 1 if spin == 0 else 2
    ai = ainfo.ptr.contents
    nalm = max(int(ai.mvstart[i]) + ai.lmax * int(ai.stride) + 1 for i in range(ai.nm))
    alms = [torch.tensor(rng.normal(size=nalm) + 1j * rng.normal(size=nalm), device=dev)
            for _ in range(nt * na)]
    maps = [torch.tensor(rng.normal(size=geom.npix), device=dev) for _ in range(nt * nm)]
    return alms, maps


@pytest.mark.parametrize("jt,spin,form", CASES)
def test_noalloc_matches_execute_device(gpu, jt, spin, form):
    rng = np.random.default_rng(11 + jt + 7 * spin + form % 97)
    synth = jt in (S.ALM2MAP, S.WY, S.ALM2MAP_DERIV1)
    for nside, lmax, nt in SIZES:
        print("noalloc case", jt, spin, form, nside, lmax, nt, flush=True)
        geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
        S.prepare(geom, ainfo, spin)
        alms, maps = _inputs(gpu, rng, geom, ainfo, jt, spin, nt)
        a1, m1 = [x.clone() for x in alms], [x.clone() for x in maps]
        a2, m2 = [x.clone() for x in alms], [x.clone() for x in maps]
        flags = S.DP | form
        op1 = S.execute_device(jt, spin, a1, m1, geom, ainfo, nt, flags)
        op2 = S.execute_device_noalloc(jt, spin, a2, m2, geom, ainfo, nt, flags)
        torch.cuda.synchronize()
        assert op1 == op2 > 0
        for got, want in (zip(m2, m1) if synth else zip(a2, a1)):
            err = rel_l2(got.cpu().numpy(), want.cpu().numpy())
            assert err <= 1e-12, (nside, lmax, nt, err)


def test_noalloc_add_and_float(gpu):
    """SHARP_ADD and single precision through the scratch area."""
    rng = np.random.default_rng(5)
    geom, ainfo = S.Geometry.healpix(16), S.AlmInfo.triangular(24)
    S.prepare(geom, ainfo, 2)
    alms, maps = _inputs(gpu, rng, geom, ainfo, S.ALM2MAP, 2, 2)
    alms = [a.to(torch.complex64) for a in alms]
    maps = [m.to(torch.float32) for m in maps]
    for jt in (S.ALM2MAP, S.MAP2ALM):
        a1, m1 = [x.clone() for x in alms], [x.clone() for x in maps]
        a2, m2 = [x.clone() for x in alms], [x.clone() for x in maps]
        S.execute_device(jt, 2, a1, m1, geom, ainfo, 2, S.ADD)
        S.execute_device_noalloc(jt, 2, a2, m2, geom, ainfo, 2, S.ADD)
        torch.cuda.synchronize()
        for got, want in (zip(m2, m1) if jt == S.ALM2MAP else zip(a2, a1)):
            assert torch.equal(got, want)

"""Device-resident entry points (include/sharp_hip.h): stream semantics and
the profiling interface.

NULL is the null stream: work on PyTorch's default stream (whose handle is
0) before and after a call is ordered with it, with no explicit
synchronisation.  A fill of the outputs with NaN right before the call, and
a reduction right after it, must both see the transform in order."""
import numpy as np
import pytest
import torch

import sharp_hip as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spin", [0, 2])
def test_default_stream_ordering(gpu, spin):
    nside, lmax = 256, 512
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    nc = 1 if spin == 0 else 2
    rng = np.random.default_rng(5)
    alms = [torch.tensor(rng.normal(size=nalm) + 1j * rng.normal(size=nalm), device=gpu)
            for _ in range(nc)]
    ref = [torch.zeros(geom.npix, dtype=torch.float64, device=gpu) for _ in range(nc)]
    S.execute_device(S.ALM2MAP, spin, alms, ref, geom, ainfo, 1, S.DP)
    torch.cuda.synchronize()
    want = [float(r.sum()) for r in ref]
    for _ in range(3):
        out = [torch.full((geom.npix,), float("nan"), dtype=torch.float64, device=gpu)
               for _ in range(nc)]
        S.execute_device(S.ALM2MAP, spin, alms, out, geom, ainfo, 1, S.DP)
        got = [float(o.sum()) for o in out]   # queued after the transform
        assert got == want, (got, want)


def test_profile_stats(gpu):
    geom, ainfo = S.Geometry.healpix(32), S.AlmInfo.triangular(64)
    nalm = S.nalm_triangular(64)
    alm = [torch.ones(nalm, dtype=torch.complex128, device=gpu)]
    m = [torch.zeros(geom.npix, dtype=torch.float64, device=gpu)]
    S.profile(True)
    S.execute_device(S.ALM2MAP, 0, alm, m, geom, ainfo, 1, S.DP)
    S.execute_device(S.MAP2ALM, 0, alm, m, geom, ainfo, 1, S.DP)
    stats = {st["name"]: st for st in S.profile_stats()}
    S.profile(False)
    assert "legendre_synth_spin0" in stats and "legendre_anal_spin0" in stats
    assert stats["legendre_synth_spin0"]["flops"] > 0
    assert all(st["ms"] >= 0 and st["launches"] >= 1 for st in stats.values())


def test_concurrent_streams_match_one_stream(gpu):
    """The library keeps work buffers per stream: T (spin 0) on one stream and
    Q/U (spin 2) on another, enqueued interleaved, give bit-identical results
    to the same calls on one stream, in both forms."""
    nside, lmax, nt = 32, 64, 3
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    nalm = S.nalm_triangular(lmax)
    g = torch.Generator(device=gpu).manual_seed(11)

    def calm():
        return torch.complex(torch.randn(nalm, generator=g, device=gpu, dtype=torch.float64),
                             torch.randn(nalm, generator=g, device=gpu, dtype=torch.float64))
    T = [calm() for _ in range(nt)]
    EB = [calm() for _ in range(2 * nt)]
    for flags in (S.DP, S.DP | S.MIXED_BF16X3):
        outs = []
        for two in (False, True):
            cur = torch.cuda.current_stream()
            sa, sb = (torch.cuda.Stream(), torch.cuda.Stream()) if two else (cur, cur)
            sa.wait_stream(cur)
            sb.wait_stream(cur)
            mt = [torch.zeros(geom.npix, dtype=torch.float64, device=gpu) for _ in range(nt)]
            mq = [torch.zeros(geom.npix, dtype=torch.float64, device=gpu) for _ in range(2 * nt)]
            at = [torch.zeros_like(T[0]) for _ in range(nt)]
            ae = [torch.zeros_like(T[0]) for _ in range(2 * nt)]
            for _ in range(2):
                S.execute_device(S.ALM2MAP, 0, T, mt, geom, ainfo, nt, flags, sa)
                S.execute_device(S.ALM2MAP, 2, EB, mq, geom, ainfo, nt, flags, sb)
                S.execute_device(S.MAP2ALM, 0, at, mt, geom, ainfo, nt, flags, sa)
                S.execute_device(S.MAP2ALM, 2, ae, mq, geom, ainfo, nt, flags, sb)
            torch.cuda.synchronize()
            outs.append([x.clone() for x in mt + mq + at + ae])
        for x, y in zip(*outs):
            assert torch.equal(x, y)

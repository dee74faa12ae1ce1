"""Batching form: transforms of uneven size in one batch (include/sharp_hip.h,
sharp_hip_batch_*).  Every job of a batch is compared with the same job run
alone through the FP64 device path: the mixed batch within 1e-6 relative L2
per output array, the FP64 batch within 1e-14.  The graph is replayed three
times with new inputs written by ordinary eager work between the replays,
and every replay is checked."""
import numpy as np
import pytest
import torch

import sharp_hip as S
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _jobs(dev, synth, seed):
    """HEALPix 16 / 32 / 8 and a Gauss grid, lmax 2*nside / 40, spins 0, 2 and 1,
    1 to 9 transforms (more than one tile group)."""
    specs = [(S.Geometry.healpix(16), S.AlmInfo.triangular(32), 0, 9),
             (S.Geometry.healpix(32), S.AlmInfo.triangular(64), 2, 3),
             (S.Geometry.iso("gauss", 41, 81), S.AlmInfo.triangular(40), 0, 2),
             (S.Geometry.healpix(8), S.AlmInfo.triangular(20, 12), 1, 1)]
    rng = np.random.default_rng(seed)
    jobs = []
    for geom, ainfo, spin, nt in specs:
        na = 1 if spin == 0 else 2
        ai = ainfo.ptr.contents
        nalm = max(int(ai.mvstart[i]) + ai.lmax * int(ai.stride) + 1 for i in range(ai.nm))
        alms = [torch.tensor(rng.normal(size=nalm) + 1j * rng.normal(size=nalm), device=dev)
                for _ in range(nt * na)]
        maps = [torch.tensor(rng.normal(size=geom.npix), device=dev) for _ in range(nt * na)]
        jt = S.ALM2MAP if synth else S.MAP2ALM
        jobs.append((jt, spin, alms, maps, geom, ainfo, nt))
    return jobs


def _alone(job, flags):
    jt, spin, alms, maps, geom, ainfo, nt = job
    a2 = [a.clone() for a in alms]
    m2 = [m.clone() for m in maps]
    S.execute_device(jt, spin, a2, m2, geom, ainfo, nt, flags)
    return a2, m2


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("synth", [True, False])
@pytest.mark.parametrize("mixed", [True, False])
def test_batch_matches_single_jobs(gpu, synth, mixed, graph):
    jobs = _jobs(gpu, synth, 3)
    flags = S.DP | (S.MIXED_FP32 if mixed else 0)
    b = S.Batch(jobs, flags, use_graph=graph)
    cfg = b.config()
    assert cfg["graph"] == graph
    rng = np.random.default_rng(9)
    for rep in range(3):
        # new inputs, written by eager work between replays
        for jt, spin, alms, maps, geom, ainfo, nt in jobs:
            src = alms if synth else maps
            for x in src:
                x.copy_(torch.tensor(rng.normal(size=x.shape) * (1 + rep)
                                     + (1j * rng.normal(size=x.shape) if x.is_complex() else 0),
                                     device=gpu, dtype=x.dtype))
        want = [_alone(j, S.DP) for j in jobs]
        torch.cuda.synchronize()
        b.execute()
        torch.cuda.synchronize()
        tol = 1e-6 if mixed else 1e-14
        for (jt, spin, alms, maps, geom, ainfo, nt), (wa, wm) in zip(jobs, want):
            outs = zip(maps, wm) if synth else zip(alms, wa)
            for got, ref in outs:
                err = rel_l2(got.cpu().numpy(), ref.cpu().numpy())
                assert err < tol, (rep, spin, nt, err)
    b.close()


def test_batch_rejects_mixed_directions(gpu):
    import subprocess, sys, textwrap, os
    code = textwrap.dedent("""
        import sys, torch
        sys.path.insert(0, %r)
        import sharp_hip as S
        g, a = S.Geometry.healpix(4), S.AlmInfo.triangular(8)
        al = [torch.zeros(45, dtype=torch.complex128, device="cuda")]
        mp = [torch.zeros(192, dtype=torch.float64, device="cuda")]
        S.Batch([(S.ALM2MAP, 0, al, mp, g, a, 1), (S.MAP2ALM, 0, al, mp, g, a, 1)], S.DP)
        """ % os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "planck-levels_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "all be syntheses or all analyses" in r.stderr, r.stderr

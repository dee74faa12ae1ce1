"""Batching form: transforms of uneven size in one batch (include/sharp_hip.h,
sharp_hip_batch_*).  Every output array of a batch is compared with the C
oracle run on the same inputs: the mixed batch within 1e-6 relative L2, the
FP64 batch within 1e-12.  The graph is replayed three times with new inputs
written by ordinary eager work between the replays, and every replay is
checked."""
import numpy as np
import pytest
import torch

import oracle as O
import sharp_hip as S
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FORMS = {"fp64": 0, "fp32": S.MIXED_FP32, "bf16x3": S.MIXED_BF16X3}
TOL = {"fp64": 1e-12, "fp32": 1e-6, "bf16x3": 1e-6}


def _spec_job(dev, rng, spec, synth):
    gkind, gpar, lmax, mmax, spin, nt = spec
    geom = S.Geometry.healpix(gpar) if gkind == "healpix" else S.Geometry.iso(gkind, *gpar)
    ainfo = S.AlmInfo.triangular(lmax, mmax)
    na = 1 if spin == 0 else 2
    ai = ainfo.ptr.contents
    nalm = max(int(ai.mvstart[i]) + ai.lmax * int(ai.stride) + 1 for i in range(ai.nm))
    alms = [torch.tensor(rng.normal(size=nalm) + 1j * rng.normal(size=nalm), device=dev)
            for _ in range(nt * na)]
    maps = [torch.tensor(rng.normal(size=geom.npix), device=dev) for _ in range(nt * na)]
    jt = S.ALM2MAP if synth else S.MAP2ALM
    return (jt, spin, alms, maps, geom, ainfo, nt), spec


# HEALPix 16 / 32 / 8 and a Gauss grid, lmax 2*nside / 40, spins 0, 2 and 1,
# 1 to 9 transforms (more than one transform group)
UNEVEN = [("healpix", 16, 32, 32, 0, 9), ("healpix", 32, 64, 64, 2, 3),
          ("gauss", (41, 81), 40, 40, 0, 2), ("healpix", 8, 20, 12, 1, 1)]


def _oracle_outputs(orc, job, spec, synth):
    """The oracle's outputs for the job's current inputs (and current output
    contents, which the parts a transform leaves alone keep)."""
    jt, spin, alms, maps, geom, ainfo, nt = job
    gkind, gpar, lmax, mmax = spec[:4]
    og = orc.healpix(gpar) if gkind == "healpix" else orc.iso(gkind, *gpar)
    oa = orc.triangular(lmax, mmax)
    a = [x.cpu().numpy().copy() for x in alms]
    m = [x.cpu().numpy().copy() for x in maps]
    orc.execute(O.ALM2MAP if synth else O.MAP2ALM, spin, a, m, og, oa, nt, O.DP)
    orc.destroy_geom(og)
    orc.destroy_alm(oa)
    return m if synth else a


def _refresh(jobs, synth, rng, scale, dev):
    for (jt, spin, alms, maps, geom, ainfo, nt), _ in jobs:
        for x in (alms if synth else maps):
            v = rng.normal(size=x.shape) * scale
            if x.is_complex():
                v = v + 1j * rng.normal(size=x.shape)
            x.copy_(torch.tensor(v, device=dev, dtype=x.dtype))


def _check_batch(orc, dev, specs, synth, form, graph, reps=3, after_create=None):
    rng = np.random.default_rng(3)
    jobs = [_spec_job(dev, rng, sp, synth) for sp in specs]
    b = S.Batch([j for j, _ in jobs], S.DP | FORMS[form], use_graph=graph)
    cfg = b.config()
    assert cfg["graph"] == graph
    assert (cfg["m_block"] in (4, 8) and cfg["ring_waves"] == 0) if synth else \
        (cfg["m_block"] == 0 and cfg["ring_waves"] in (2, 4)), cfg
    if after_create is not None:
        after_create()
    for rep in range(reps):
        _refresh(jobs, synth, rng, 1 + rep, dev)
        torch.cuda.synchronize()
        want = [_oracle_outputs(orc, j, sp, synth) for j, sp in jobs]
        b.execute()
        torch.cuda.synchronize()
        for ((jt, spin, alms, maps, geom, ainfo, nt), sp), w in zip(jobs, want):
            got = maps if synth else alms
            for g, r in zip(got, w):
                err = rel_l2(g.cpu().numpy(), r)
                assert err < TOL[form], (rep, sp, form, err)
                if form != "fp64":
                    assert err > 1e-13, (rep, sp, "matrix-core path did not run", err)
    b.close()


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("synth", [True, False])
@pytest.mark.parametrize("form", sorted(FORMS))
def test_batch_vs_oracle(gpu, oracle, synth, form, graph):
    _check_batch(oracle, gpu, UNEVEN, synth, form, graph)


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("form", sorted(FORMS))
def test_batch_one_geometry_several_spins(gpu, oracle, form, graph):
    """Analysis jobs of spins 0, 2 and 1 on one geometry (a Nside no other
    test uses, so its cache entry is new): each job's per-ring m limits stay
    valid while the later jobs add theirs."""
    nside = {("fp64", True): 12, ("fp64", False): 14, ("fp32", True): 18,
             ("fp32", False): 20, ("bf16x3", True): 22, ("bf16x3", False): 26}[(form, graph)]
    specs = [("healpix", nside, 30, 30, 0, 2), ("healpix", nside, 30, 30, 2, 1),
             ("healpix", nside, 30, 30, 1, 1), ("healpix", nside, 26, 20, 2, 1)]
    _check_batch(oracle, gpu, specs, False, form, graph, reps=2)


def test_batch_survives_release(gpu, oracle):
    """sharp_hip_release with a live batch: the batch keeps the tables of its
    jobs and still computes the right answer afterwards."""
    _check_batch(oracle, gpu, UNEVEN[:2], False, "fp32", True, reps=2, after_create=S.release)
    _check_batch(oracle, gpu, UNEVEN[:2], True, "fp64", False, reps=2, after_create=S.release)
    _check_batch(oracle, gpu, UNEVEN, True, "bf16x3", True, reps=2, after_create=S.release)


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


def test_tuner_cache_file(gpu, tmp_path):
    """SHARP_HIP_TUNE_CACHE: the first process tunes and appends its choice
    (key, transform group, workgroups per CU, tile shape, flat); a second process
    finds it there and uses the same tile shape without tuning again."""
    import json, os, subprocess, sys, textwrap
    cache = tmp_path / "tune.txt"
    code = textwrap.dedent("""
        import json, sys, torch
        sys.path.insert(0, %r)
        import sharp_hip as S
        g, a = S.Geometry.healpix(16), S.AlmInfo.triangular(32)
        al = [torch.zeros(561, dtype=torch.complex128, device="cuda") for _ in range(3)]
        mp = [torch.zeros(3072, dtype=torch.float64, device="cuda") for _ in range(3)]
        b = S.Batch([(S.MAP2ALM, 0, al, mp, g, a, 3)], S.DP | S.MIXED_FP32)
        print(json.dumps(b.config()))
        """ % os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "planck-levels_amd"))
    env = dict(os.environ, SHARP_HIP_TUNE_CACHE=str(cache))
    outs = []
    for _ in range(2):
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           timeout=110, env=env)
        assert r.returncode == 0, r.stderr[-2000:]
        outs.append(json.loads(r.stdout.strip().splitlines()[-1]))
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and len(lines[0].split()) == 5, lines
    assert outs[0] == outs[1]

"""Parity at the full sizes of the workloads this engine is measured on,
each checked against the reference build (oracle/_ref, libsharp's own FP64
sharp_execute with OpenMP) with the error metric of
libsharp/sharp_testsuite.c:193-229 (per component, relative L2):

* HEALPix Nside 1024, lmax 2048, TQU, FP64 (spectra C^TT = 1/(l(l+1)+1),
  EE = 0.1 TT, BB = 0.01 TT, TE correlation 0.5): alm2map and map2alm each
  within 1e-12 of the reference;
* HEALPix Nside 2048, lmax 4096, TQU, 8 realisations (same spectra times a 5'
  beam): the BF16x3 and FP32 matrix-core forms within 1e-6 of the FP64 form
  on all 24 maps and all 24 round-trip a_lm, and realisation 0 of the FP64
  form within 1e-12 of the reference (every HEALPix ring length 4i, i < 2048,
  and 8192 goes through the ring FFT here);
* the uneven batch: 24 spin-0 maps at each of Nside 128/256/512/1024 (lmax =
  2 Nside) and 8 Gauss-Legendre grids at each of lmax 255/767, as two
  batches (synthesis, analysis) replayed from HIP graphs, in each form: on
  FP32 and on BF16-split matrix cores every map and round-trip a_lm within
  1e-6 of the FP64 form run per geometry, the FP64 batch within 1e-12 of it,
  and the FP64 form within 1e-12 of the reference on one transform per grid.

The reference runs on the host's cores (16 threads on the GPU box)."""
import os

import numpy as np
import pytest
import torch

import oracle as O
import sharp_hip as S
from sht_inputs import make_alms, spin0_alms

pytestmark = pytest.mark.gpu


def _rel(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def _trel(x, y):
    return float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y))


def _zeros_maps(n, npix, dev):
    return [torch.zeros(npix, dtype=torch.float64, device=dev) for _ in range(n)]


def _tqu_gpu(geom, ainfo, T, EB, flags, dev, npix):
    """alm2map (T; E,B -> Q,U) then map2alm of those maps, all realisations."""
    nb = len(T)
    mt, mqu = _zeros_maps(nb, npix, dev), _zeros_maps(2 * nb, npix, dev)
    S.execute_device(S.ALM2MAP, 0, T, mt, geom, ainfo, nb, flags)
    S.execute_device(S.ALM2MAP, 2, EB, mqu, geom, ainfo, nb, flags)
    at = [torch.zeros_like(T[0]) for _ in range(nb)]
    aeb = [torch.zeros_like(T[0]) for _ in range(2 * nb)]
    S.execute_device(S.MAP2ALM, 0, at, mt, geom, ainfo, nb, flags)
    S.execute_device(S.MAP2ALM, 2, aeb, mqu, geom, ainfo, nb, flags)
    torch.cuda.synchronize()
    return mt, mqu, at, aeb


def _ref_tqu(ref, nside, lmax, t, e, b, maps_for_analysis):
    """The reference's alm2map of (t, e, b) and its map2alm of the given
    (T, Q, U) maps (numpy)."""
    g, a = ref.healpix(nside), ref.triangular(lmax)
    npix = 12 * nside * nside
    mt, mqu = [np.zeros(npix)], [np.zeros(npix), np.zeros(npix)]
    ref.execute(O.ALM2MAP, 0, [t], mt, g, a, 1, O.DP)
    ref.execute(O.ALM2MAP, 2, [e, b], mqu, g, a, 1, O.DP)
    at, aeb = [np.zeros_like(t)], [np.zeros_like(t), np.zeros_like(t)]
    ref.execute(O.MAP2ALM, 0, at, [maps_for_analysis[0]], g, a, 1, O.DP)
    ref.execute(O.MAP2ALM, 2, aeb, list(maps_for_analysis[1:]), g, a, 1, O.DP)
    ref.destroy_geom(g)
    ref.destroy_alm(a)
    return (mt[0], mqu[0], mqu[1]), (at[0], aeb[0], aeb[1])


def _check_vs_ref(ref, nside, lmax, T, E, B, mt, mqu, at, aeb, i=0):
    nb = len(T)
    gm = (mt[i].cpu().numpy(), mqu[2 * i].cpu().numpy(), mqu[2 * i + 1].cpu().numpy())
    ga = (at[i].cpu().numpy(), aeb[2 * i].cpu().numpy(), aeb[2 * i + 1].cpu().numpy())
    rm, ra = _ref_tqu(ref, nside, lmax, T[i].cpu().numpy(), E[i].cpu().numpy(),
                      B[i].cpu().numpy(), gm)
    errs = [_rel(x, y) for x, y in zip(gm, rm)] + [_rel(x, y) for x, y in zip(ga, ra)]
    assert max(errs) <= 1e-12, errs
    assert nb >= 1
    return errs


def test_nside1024_lmax2048_tqu_fp64_vs_reference(gpu, ref):
    nside, lmax = 1024, 2048
    print("case: Nside 1024 lmax 2048 TQU FP64 vs reference", flush=True)
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    T, E, B = make_alms(torch, lmax, 1, 2024, gpu, fwhm_arcmin=0.0)
    EB = [E[0], B[0]]
    mt, mqu, at, aeb = _tqu_gpu(geom, ainfo, T, EB, S.DP, gpu, 12 * nside * nside)
    print("errors T Q U / T E B:", _check_vs_ref(ref, nside, lmax, T, E, B, mt, mqu, at, aeb))


def test_nside2048_lmax4096_tqu8_mixed_and_fp64(gpu, ref):
    nside, lmax, nb = 2048, 4096, 8
    npix = 12 * nside * nside
    print("case: Nside 2048 lmax 4096 TQU x8", flush=True)
    geom, ainfo = S.Geometry.healpix(nside), S.AlmInfo.triangular(lmax)
    T, E, B = make_alms(torch, lmax, nb, 1234, gpu)
    EB = [x for pair in zip(E, B) for x in pair]
    f_mt, f_mqu, f_at, f_aeb = _tqu_gpu(geom, ainfo, T, EB, S.DP, gpu, npix)
    errs = _check_vs_ref(ref, nside, lmax, T, E, B, f_mt, f_mqu, f_at, f_aeb)
    print("FP64 realisation 0 vs reference:", errs, flush=True)
    for name, form in (("bf16x3", S.MIXED_BF16X3), ("fp32", S.MIXED_FP32)):
        mt, mqu, at, aeb = _tqu_gpu(geom, ainfo, T, EB, S.DP | form, gpu, npix)
        e = [_trel(x, y) for x, y in zip(mt + mqu, f_mt + f_mqu)]
        e += [_trel(x, y) for x, y in zip(at + aeb, f_at + f_aeb)]
        print(name, "vs FP64: max %.3g min %.3g" % (max(e), min(e)), flush=True)
        assert max(e) <= 1e-6, (name, max(e))
        assert min(e) > 1e-12, (name, "matrix-core form did not run")
        del mt, mqu, at, aeb
        torch.cuda.empty_cache()


def test_uneven_batch_vs_fp64_and_reference(gpu, ref):
    print("case: uneven batch (96 HEALPix + 16 Gauss-Legendre)", flush=True)
    specs = [("healpix", n, 2 * n, 24) for n in (128, 256, 512, 1024)]
    specs += [("gauss", 255, 255, 8), ("gauss", 767, 767, 8)]
    gen = torch.Generator(device=gpu).manual_seed(77)
    jobs = []
    for kind, n, lmax, nt in specs:
        geom = (S.Geometry.healpix(n) if kind == "healpix"
                else S.Geometry.iso("gauss", lmax + 1, 2 * lmax + 1))
        ainfo = S.AlmInfo.triangular(lmax)
        alms = spin0_alms(torch, lmax, nt, gen, gpu)
        jobs.append(dict(spec=(kind, n, lmax, nt), geom=geom, ainfo=ainfo, alms=alms))
    # the FP64 form per geometry, and transform 0 of every grid against the
    # reference
    for j in jobs:
        kind, n, lmax, nt = j["spec"]
        m64 = _zeros_maps(nt, j["geom"].npix, gpu)
        S.execute_device(S.ALM2MAP, 0, j["alms"], m64, j["geom"], j["ainfo"], nt, S.DP)
        a64 = [torch.zeros_like(j["alms"][0]) for _ in range(nt)]
        S.execute_device(S.MAP2ALM, 0, a64, m64, j["geom"], j["ainfo"], nt, S.DP)
        torch.cuda.synchronize()
        j["m64"], j["a64"] = m64, a64
        og = ref.healpix(n) if kind == "healpix" else ref.iso("gauss", lmax + 1, 2 * lmax + 1)
        oa = ref.triangular(lmax)
        rm = [np.zeros(j["geom"].npix)]
        ref.execute(O.ALM2MAP, 0, [j["alms"][0].cpu().numpy()], rm, og, oa, 1, O.DP)
        ra = [np.zeros(j["alms"][0].numel(), np.complex128)]
        ref.execute(O.MAP2ALM, 0, ra, [m64[0].cpu().numpy()], og, oa, 1, O.DP)
        ref.destroy_geom(og)
        ref.destroy_alm(oa)
        er = (_rel(m64[0].cpu().numpy(), rm[0]), _rel(a64[0].cpu().numpy(), ra[0]))
        assert max(er) <= 1e-12, (j["spec"], er)
    # every form as two batches (synthesis, analysis) replayed from graphs
    for form, flag, tol in (("fp32", S.MIXED_FP32, 1e-6), ("bf16x3", S.MIXED_BF16X3, 1e-6),
                            ("fp64", 0, 1e-12)):
        for j in jobs:
            j["maps"] = _zeros_maps(j["spec"][3], j["geom"].npix, gpu)
            j["out"] = [torch.zeros_like(j["alms"][0]) for _ in range(j["spec"][3])]
        bs = S.Batch([(S.ALM2MAP, 0, j["alms"], j["maps"], j["geom"], j["ainfo"], j["spec"][3])
                      for j in jobs], S.DP | flag, use_graph=True)
        ba = S.Batch([(S.MAP2ALM, 0, j["out"], j["maps"], j["geom"], j["ainfo"], j["spec"][3])
                      for j in jobs], S.DP | flag, use_graph=True)
        assert bs.config()["graph"] and ba.config()["graph"]
        for rep in range(2):
            bs.execute()
            ba.execute()
            torch.cuda.synchronize()
        worst = 0.0
        for j in jobs:
            e = [_trel(x, y) for x, y in zip(j["maps"], j["m64"])]
            e += [_trel(x, y) for x, y in zip(j["out"], j["a64"])]
            assert max(e) <= tol, (form, j["spec"], max(e))
            if form != "fp64":
                assert min(e) > 1e-12, (form, j["spec"], "matrix-core form did not run")
            worst = max(worst, max(e))
        print(form, "batch vs FP64 per geometry: worst %.3g" % worst, bs.config(), ba.config(),
              flush=True)
        bs.close()
        ba.close()

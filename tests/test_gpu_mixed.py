"""Mixed-precision forms: FP64 recurrence with exponent scaling, contraction
on the matrix cores -- FP32 MFMA (SHARP_HIP_MIXED_FP32) or BF16 MFMA on exact
three-piece BF16 splits of the FP32 operands (SHARP_HIP_MIXED_BF16X3).  The
outputs are compared with the FP64 oracle (small cases) or the FP64 GPU path
(larger cases, itself pinned to the oracle and the reference build at 1e-12).

Tolerance: 1e-6 relative L2 per output array, the bound the mixed forms are
specified to meet.  Every case also has to be far from FP64-exact (>1e-12),
which shows that the MFMA path, not the FP64 one, produced it.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from sht_cases import healpix, iso, rect, run_case, tri
from test_gpu_parity import _general_geom

pytestmark = pytest.mark.gpu

MIXED_FP32 = 1 << 24     # SHARP_HIP_MIXED_FP32 (include/sharp_hip.h)
MIXED_BF16X3 = 1 << 25   # SHARP_HIP_MIXED_BF16X3
TOLMIX = 1e-6


@pytest.fixture(params=[MIXED_FP32, MIXED_BF16X3], ids=["fp32", "bf16x3"])
def form(request):
    return request.param


def compare_mixed(product, checker, *args, flags=O.DP, tol=TOLMIX, form=MIXED_FP32, **kw):
    pa, pm = run_case(product, *args, flags | form, **kw)
    ca, cm = run_case(checker, *args, flags, **kw)
    errs = [rel_l2(x, y) for x, y in zip(pa + pm, ca + cm)]
    print("mixed rel err", args[2:], max(errs))
    assert max(errs) < tol, errs
    # FP32-level differences show that the matrix-core path produced the result
    assert max(errs) > 1e-12, errs
    return max(errs)


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
@pytest.mark.parametrize("ntrans", [1, 3, 8])
def test_healpix(product, oracle, gpu, spin, jobtype, ntrans, form):
    compare_mixed(product, oracle, healpix(16), tri(32), jobtype, spin, ntrans, form=form)


@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM, O.Yt, O.WY])
@pytest.mark.parametrize("spin", [0, 1, 2, 3])
def test_gauss_jobtypes(product, oracle, gpu, jobtype, spin, form):
    compare_mixed(product, oracle, iso("gauss", 25, 49), tri(24), jobtype, spin, 2, form=form)


@pytest.mark.parametrize("spin,ntrans", [(0, 16), (0, 17), (2, 9), (1, 11)])
def test_several_launches(product, oracle, gpu, spin, ntrans, form):
    """More transforms than one launch holds (16 for spin 0, 8 for spin s)."""
    for jt in (O.ALM2MAP, O.MAP2ALM):
        compare_mixed(product, oracle, iso("gauss", 17, 33), tri(16), jt, spin, ntrans, form=form)


@pytest.mark.parametrize("spin", [0, 1, 2, 3])
def test_band_limit_near_spin(product, oracle, gpu, spin, form):
    """lmax = spin .. spin+3: rows whose recurrence has no step, one step (the
    first, whose lambda_{l-2} term is zero), or the first steps with a scale
    factor s_{m,l} != 1 in the unit-coefficient tables of the mixed forms."""
    for lmax in range(spin, spin + 4):
        for jt in (O.ALM2MAP, O.MAP2ALM):
            compare_mixed(product, oracle, iso("gauss", 9, 17), tri(lmax), jt, spin, 2,
                          form=form)


def test_deriv1(product, oracle, gpu, form):
    compare_mixed(product, oracle, healpix(8), tri(20), O.DERIV1, 1, 2, form=form)


@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_flags_and_layouts(product, oracle, gpu, jobtype, form):
    compare_mixed(product, oracle, healpix(8), tri(18), jobtype, 2, 2,
                  flags=O.DP | O.ADD | O.REAL_HARMONICS, form=form)
    compare_mixed(product, oracle, healpix(8), rect(20, 12), jobtype, 0, 3, form=form)
    compare_mixed(product, oracle, healpix(8), tri(20, 7), jobtype, 2, 1, form=form)
    compare_mixed(product, oracle, healpix(8), tri(16), jobtype, 0, 2, flags=0,
                  tol=2e-6, form=form)


@pytest.mark.parametrize("spin", [0, 1, 2])
@pytest.mark.parametrize("jobtype", [O.ALM2MAP, O.MAP2ALM])
def test_general_geometry(product, oracle, gpu, spin, jobtype, form):
    compare_mixed(product, oracle, _general_geom, tri(10), jobtype, spin, 2, form=form)


def test_high_spin(product, oracle, gpu, form):
    """spin 30: rows l < spin carry no lambda and stay zero."""
    for jt in (O.ALM2MAP, O.MAP2ALM):
        compare_mixed(product, oracle, iso("gauss", 64, 127), tri(63), jt, 30, 1, form=form)


@pytest.mark.parametrize("spin", [0, 2])
def test_larger_vs_fp64(product, gpu, spin, form):
    """HEALPix nside 128, lmax 256, 8 transforms, against the FP64 GPU path;
    both directions, and far from FP64-exact."""
    for jt in (O.ALM2MAP, O.MAP2ALM):
        pa, pm = run_case(product, healpix(128), tri(256), jt, spin, 8, O.DP | form)
        ca, cm = run_case(product, healpix(128), tri(256), jt, spin, 8, O.DP)
        outs = zip(pm, cm) if jt == O.ALM2MAP else zip(pa, ca)
        errs = [rel_l2(x, y) for x, y in outs]
        print("mixed vs fp64", jt, spin, max(errs), min(errs))
        assert max(errs) < TOLMIX, errs
        assert min(errs) > 1e-12, errs


def _packed96(lib):
    """Packed layout (real-only m = 0 block), lmax 96."""
    lmax = 96
    mval = np.arange(lmax + 1)
    mvs = np.zeros(lmax + 1, np.int64)
    pos = lmax + 1
    for m in range(1, lmax + 1):
        mvs[m] = pos - 2 * m
        pos += 2 * (lmax - m + 1)
    return lib.general_alm(lmax, mval, mvs, 1, O.PACKED)


@pytest.mark.parametrize("case", ["second_launch", "deriv1", "packed_add"])
def test_whole_groups(product, oracle, gpu, case, form):
    """lmax 96, so that the unscaled part has whole 32-l groups (the BF16
    synthesis reads its a_lm pre-split there): transforms 8 and up (a second
    launch, its first column past the first 32), DERIV1 (its own column
    pattern), and a packed a_lm layout with SHARP_ADD through the gather."""
    if case == "second_launch":
        for jt in (O.ALM2MAP, O.MAP2ALM):
            compare_mixed(product, oracle, iso("gauss", 97, 193), tri(96), jt, 2, 10, form=form)
    elif case == "deriv1":
        compare_mixed(product, oracle, healpix(32), tri(96), O.DERIV1, 1, 3, form=form)
    else:
        for jt in (O.ALM2MAP, O.MAP2ALM):
            compare_mixed(product, oracle, healpix(32), _packed96, jt, 2, 2,
                          flags=O.DP | O.ADD, form=form)


def test_analysis_in_m_chunks(gpu, tmp_path):
    """SHARP_HIP_PART_CAP_MB caps the per-group partial sums of an analysis
    launch; above it the analysis runs in chunks of orders m (Legendre launch
    and scatter per chunk).  With a 1 MiB cap (about 20 chunks here) the
    outputs equal those of one chunk, for the BF16-split, FP32 and FP64
    forms."""
    import os, subprocess, sys, textwrap
    code = textwrap.dedent("""
        import sys, numpy as np, torch
        sys.path.insert(0, %r)
        import sharp_hip as S
        g, a = S.Geometry.healpix(256), S.AlmInfo.triangular(512)
        rng = np.random.default_rng(4)
        maps = [torch.tensor(rng.normal(size=g.npix), device="cuda") for _ in range(8)]
        out = []
        for form in (S.MIXED_BF16X3, S.MIXED_FP32, 0):
            for spin, nc in ((0, 4), (2, 8)):
                al = [torch.zeros(S.nalm_triangular(512), dtype=torch.complex128,
                                  device="cuda") for _ in range(nc)]
                S.execute_device(S.MAP2ALM, spin, al, maps[:nc], g, a, nc // (1 if spin == 0 else 2),
                                 S.DP | form)
                torch.cuda.synchronize()
                out += [x.cpu().numpy() for x in al]
        np.save(sys.argv[1], np.stack(out))
        """ % os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "planck-levels_amd"))
    res = []
    for cap in (None, "1"):
        env = dict(os.environ)
        env.pop("SHARP_HIP_PART_CAP_MB", None)
        if cap:
            env["SHARP_HIP_PART_CAP_MB"] = cap
        f = str(tmp_path / ("out%s.npy" % (cap or "")))
        r = subprocess.run([sys.executable, "-c", code, f], capture_output=True, text=True,
                           timeout=110, env=env)
        assert r.returncode == 0, r.stderr[-2000:]
        res.append(np.load(f))
    # the same partial sums added in the same order: identical
    assert np.array_equal(res[0], res[1])

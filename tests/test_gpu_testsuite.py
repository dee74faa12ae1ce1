"""The reference's own test program, libsharp/sharp_testsuite.c, compiled
unchanged and linked against the drop-in library (oracle/Makefile,
testsuite_hip), run on the GPU:

* `acctest` (libsharp/sharp_testsuite.c:503-529): check_sign_scale's golden
  values for ntrans 1..9 and spins 0, 1, 2 and DERIV1 (rel. 1e-12), then
  Gauss-Legendre round trips at lmax 127 for spins 0, 1, 2, 3, 30, ntrans
  1..6 and every nv, each within 1e-10;
* `test` (:531-589) on HEALPix grids: the printed round-trip errors equal
  those of the same program linked against the reference build."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_TS = os.path.join(ROOT, "oracle", "_ref", "sharp_testsuite_hip")
REF_TS = os.path.join(ROOT, "oracle", "_ref", "sharp_testsuite")


def _run(exe, *args, threads=None):
    if not os.path.exists(exe):
        pytest.skip("%s not built (make -C oracle ref testsuite_hip)" % os.path.basename(exe))
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    return subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=110,
                          env=env)


def test_reference_acctest_through_dropin(gpu):
    r = _run(HIP_TS, "acctest")
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert r.stdout.count("Passed.") == 2, r.stdout[-2000:]


def _errors(out):
    """(max rel, max abs) round-trip error from the result line of `test`."""
    line = [l for l in out.splitlines() if l.split()[1:2] == ["healpix"]][-1]
    f = line.split()
    return float(f[-2]), float(f[-1])


@pytest.mark.parametrize("args", [("healpix", "255", "-1", "128", "-1", "0", "3"),
                                  ("healpix", "383", "200", "160", "-1", "2", "2")])
def test_reference_test_mode_matches_reference_build(gpu, args):
    r = _run(HIP_TS, "test", *args)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    q = _run(REF_TS, "test", *args, threads=16)
    assert q.returncode == 0, q.stderr[-2000:]
    (er, ea), (qr, qa) = _errors(r.stdout), _errors(q.stdout)
    # HEALPix round trips are not exact: both show the same quadrature error
    assert 1e-8 < qr < 1e-1
    assert abs(er - qr) <= 1e-9 * qr and abs(ea - qa) <= 1e-9 * qa, (er, qr, ea, qa)

"""The drop-in boundary, checked without a GPU: exported symbols, struct
layout, the info objects, and the reference's error behaviour."""
import ctypes as ct
import os
import re
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "planck-levels_amd", "libsharp_hip.so")
REF = "/root/reference"


def _declared_functions():
    names = set()
    for h in os.listdir(INC):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(INC, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|ptrdiff_t|double)\s+\*?\s*(sharp_\w+)\s*\(",
                             txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = _declared_functions()
    # static inline helpers of sharp_geomhelpers.h live in the header
    declared -= {"sharp_make_healpix_geom_info", "sharp_make_ecp_geom_info"}
    assert declared, "no declarations parsed"
    missing = sorted(declared - exported)
    assert not missing, missing
    # the libsharp path symbols (libsharp/sharp_lowlevel.h, geom/alm helpers)
    for s in ("sharp_execute", "sharp_make_geom_info", "sharp_destroy_geom_info",
              "sharp_make_alm_info", "sharp_make_general_alm_info", "sharp_alm_index",
              "sharp_destroy_alm_info", "sharp_set_chunksize_min",
              "sharp_set_nchunks_max", "sharp_make_weighted_healpix_geom_info",
              "sharp_make_gauss_geom_info", "sharp_make_fejer1_geom_info",
              "sharp_make_cc_geom_info", "sharp_make_fejer2_geom_info",
              "sharp_make_mw_geom_info", "sharp_make_triangular_alm_info",
              "sharp_make_rectangular_alm_info", "sharp_get_mlim"):
        assert s in exported, s


def test_struct_layout(tmp_path):
    """sizeof/offsetof of the ABI structs as seen by a C caller of include/."""
    src = tmp_path / "layout.c"
    src.write_text(textwrap.dedent("""
        #include <stdio.h>
        #include <stddef.h>
        #include "sharp_lowlevel.h"
        int main(void) {
          printf("%zu %zu %zu %zu %zu\\n", sizeof(sharp_ringinfo),
                 offsetof(sharp_ringinfo, ofs), offsetof(sharp_ringinfo, nph),
                 sizeof(sharp_ringpair), sizeof(sharp_geom_info));
          printf("%zu %zu %zu %zu\\n", sizeof(sharp_alm_info),
                 offsetof(sharp_alm_info, mval), offsetof(sharp_alm_info, mvstart),
                 offsetof(sharp_alm_info, stride));
          printf("%d %d %d %d %d\\n", SHARP_DP, SHARP_ADD, SHARP_REAL_HARMONICS,
                 SHARP_NO_FFT, SHARP_ALM2MAP_DERIV1);
          return 0; }
        """))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", INC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    l1, l2, l3 = out.split("\n")[:3]
    assert list(map(int, l1.split())) == [56, 40, 48, 112, 16]
    assert list(map(int, l2.split())) == [40, 8, 24, 32]
    assert list(map(int, l3.split())) == [16, 32, 64, 128, 4]
    assert ct.sizeof(O.RingInfo) == 56 and ct.sizeof(O.AlmInfo) == 40


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "libsharp")),
                    reason="reference sources not present")
def test_reference_headers_agree(tmp_path):
    """A translation unit compiled against the reference's headers sees the
    same layout and enum values (the drop-in is link-compatible)."""
    src = tmp_path / "ref_layout.c"
    src.write_text(textwrap.dedent("""
        #include <stdio.h>
        #include <stddef.h>
        #include "sharp_lowlevel.h"
        int main(void) {
          printf("%zu %zu %zu %zu %zu\\n", sizeof(sharp_ringinfo),
                 offsetof(sharp_ringinfo, ofs), offsetof(sharp_ringinfo, nph),
                 sizeof(sharp_ringpair), sizeof(sharp_geom_info));
          printf("%zu %zu %zu %zu\\n", sizeof(sharp_alm_info),
                 offsetof(sharp_alm_info, mval), offsetof(sharp_alm_info, mvstart),
                 offsetof(sharp_alm_info, stride));
          printf("%d %d %d %d %d\\n", SHARP_DP, SHARP_ADD, SHARP_REAL_HARMONICS,
                 SHARP_NO_FFT, SHARP_ALM2MAP_DERIV1);
          return 0; }
        """))
    exe = tmp_path / "ref_layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(REF, "libsharp"), str(src),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    l1, l2, l3 = out.split("\n")[:3]
    assert list(map(int, l1.split())) == [56, 40, 48, 112, 16]
    assert list(map(int, l2.split())) == [40, 8, 24, 32]
    assert list(map(int, l3.split())) == [16, 32, 64, 128, 4]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "libsharp")),
                    reason="reference sources not present")
def test_sharp_cxx_wrapper_links(tmp_path):
    """The reference's C++ wrapper (libsharp/sharp_cxx.h) compiles unchanged
    against include/ and links against the drop-in library."""
    src = tmp_path / "cxx.cc"
    src.write_text(textwrap.dedent("""
        #include "sharp_cxx.h"
        int main(int argc, char **) {
          sharp_cxxjob<double> job;
          job.set_Healpix_geometry(4);
          job.set_triangular_alm_info(8, 8);
          if (argc > 5) {   // not executed: no GPU needed to link
            double alm[90] = {0}, map[192] = {0};
            job.alm2map(alm, map, false);
            job.map2alm(map, alm, true);
          }
          return 0; }
        """))
    # the wrapper is placed next to the test source in the temporary directory,
    # so its own #include "sharp_lowlevel.h" etc. resolve to include/
    import shutil
    shutil.copy(os.path.join(REF, "libsharp", "sharp_cxx.h"), tmp_path / "sharp_cxx.h")
    exe = tmp_path / "cxx"
    subprocess.run(["g++", "-I", str(tmp_path), "-I", INC, str(src), "-o", str(exe), LIB,
                    "-Wl,-rpath," + os.path.dirname(LIB)], check=True)
    subprocess.run([str(exe)], check=True)


def _run_py(code):
    """Runs code in a fresh interpreter with the library loaded (ctypes)."""
    prog = textwrap.dedent("""
        import ctypes as ct, sys
        sys.path.insert(0, %r)
        import oracle as O
        L = O.SharpLike(%r, "sharp_")
        """ % (os.path.join(ROOT, "oracle"), LIB)) + textwrap.dedent(code)
    return subprocess.run([sys.executable, "-c", prog], capture_output=True, text=True,
                          timeout=120)


@pytest.mark.parametrize("code,msg", [
    ("g=L.healpix(4); a=L.triangular(8); L.execute(1,0,[],[],g,a,0,O.DP)",
     "bad number of simultaneous transforms"),
    ("g=L.healpix(4); a=L.triangular(8); L.execute(1,0,[],[],g,a,101,O.DP)",
     "bad number of simultaneous transforms"),
    ("g=L.healpix(4); a=L.triangular(8); L.execute(1,9,[],[],g,a,1,O.DP)", "bad spin"),
    ("g=L.healpix(4); a=L.general_alm(8,[0,1,3],[0,9,20]); L.execute(1,0,[],[],g,a,1,O.DP)",
     "not all m values are present"),
    ("g=L.healpix(4); a=L.general_alm(8,[0,1,1],[0,9,20]); L.execute(1,0,[],[],g,a,1,O.DP)",
     "duplicate m value"),
    ("g=L.iso('gauss',5,7); a=L.triangular(8); L.execute(1,0,[],[],g,a,1,O.DP|O.NO_FFT)",
     "bad ring size"),
    ("a=L.general_alm(4,[0,1,2,3,4],[0,5,9,12,14],1,O.PACKED); L.lib.sharp_alm_index(a,2,1)",
     "sharp_alm_index not applicable with SHARP_PACKED alms"),
])
def test_error_behaviour(code, msg):
    """Invalid arguments print the reference's message and exit with status 1
    (c_utils/c_utils.c:35-39), before any GPU work."""
    r = _run_py(code)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert msg in r.stderr, r.stderr


def test_noalloc_rejects_oversize_job():
    """sharp_hip_execute_device_noalloc refuses a job whose work memory does
    not fit its fixed scratch area, with the sizes, before any device call."""
    r = _run_py("""
        import ctypes as ct
        g=L.healpix(1024); a=L.triangular(2048)
        f=L.lib.sharp_hip_execute_device_noalloc
        f.argtypes=[ct.c_int,ct.c_int,ct.c_void_p,ct.c_void_p,ct.POINTER(O.GeomInfo),
                    ct.POINTER(O.AlmInfo),ct.c_int,ct.c_int,ct.c_void_p,ct.c_void_p]
        p=(ct.c_void_p*1)()
        f(1,0,p,p,g,a,1,O.DP,None,None)
        """)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "more than the %d of the fixed scratch area" % (64 << 20) in r.stderr, r.stderr


def test_fails_loudly_without_gpu():
    """No CPU fallback: a valid transform on a host without a GPU stops with a
    message instead of computing anything."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    r = _run_py("""
        import numpy as np
        g=L.healpix(4); a=L.triangular(8)
        L.execute(1,0,[np.zeros(45,complex)],[np.zeros(192)],g,a,1,O.DP)
        """)
    assert r.returncode == 1 and "no GPU device available" in r.stderr, r.stderr


@pytest.mark.parametrize("maker", [
    lambda L: L.healpix(32), lambda L: L.healpix(5, np.linspace(0.5, 1.5, 10)),
    lambda L: L.iso("gauss", 65, 128), lambda L: L.iso("fejer1", 20, 41),
    lambda L: L.iso("cc", 21, 40), lambda L: L.iso("fejer2", 19, 38),
    lambda L: L.iso("mw", 17, 33)])
def test_geometry_objects_match_reference(product, ref, maker):
    from test_oracle import _geom_eq
    _geom_eq(maker(product), maker(ref))


def test_alm_objects_and_mlim_match_reference(product, ref):
    for mk in (lambda L: L.triangular(30, 20), lambda L: L.rectangular(30, 20),
               lambda L: L.triangular(7, 7, 3)):
        a1, a2 = mk(product).contents, mk(ref).contents
        assert (a1.lmax, a1.nm, a1.flags, a1.stride) == (a2.lmax, a2.nm, a2.flags, a2.stride)
        for i in range(a1.nm):
            assert a1.mval[i] == a2.mval[i] and a1.mvstart[i] == a2.mvstart[i]
    rng = np.random.default_rng(4)
    for _ in range(200):
        th = rng.uniform(0, np.pi)
        lmax, spin = int(rng.integers(1, 9000)), int(rng.integers(0, 5))
        assert product.mlim(lmax, spin, np.sin(th), np.cos(th)) == \
            ref.mlim(lmax, spin, np.sin(th), np.cos(th))

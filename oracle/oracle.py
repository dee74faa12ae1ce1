"""ctypes loader for the CPU checkers (TEST INFRASTRUCTURE ONLY).

Two checkers live under ``oracle/``:

* ``liboracle.so`` -- the plain-C restatement in ``sht_oracle.c`` (prefix
  ``orc_``); built by ``make -C oracle oracle``.
* ``_ref/libsharp_ref.so`` -- the reference's own libsharp + libfftpack +
  c_utils compiled from their sources in place (prefix ``sharp_``); built by
  ``make -C oracle ref`` where the reference sources are present.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  The drop-in library never does.

``SharpLike`` wraps any library exporting the libsharp C interface under a
given symbol prefix, so the same code drives the restatement, the reference
build and the drop-in library (``sharp_`` prefix as well).
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsharp_ref.so")

# job types / flags: libsharp/sharp_lowlevel.h:162-191
MAP2ALM, ALM2MAP, Yt, WY, DERIV1 = 0, 1, 2, 3, 4
DP, ADD, REAL_HARMONICS, NO_FFT = 1 << 4, 1 << 5, 1 << 6, 1 << 7
PACKED = 1


class RingInfo(ct.Structure):
    _fields_ = [("theta", ct.c_double), ("phi0", ct.c_double),
                ("weight", ct.c_double), ("cth", ct.c_double),
                ("sth", ct.c_double), ("ofs", ct.c_ssize_t),
                ("nph", ct.c_int), ("stride", ct.c_int)]


class RingPair(ct.Structure):
    _fields_ = [("r1", RingInfo), ("r2", RingInfo)]


class GeomInfo(ct.Structure):
    _fields_ = [("pair", ct.POINTER(RingPair)), ("npairs", ct.c_int),
                ("nphmax", ct.c_int)]


class AlmInfo(ct.Structure):
    _fields_ = [("lmax", ct.c_int), ("nm", ct.c_int),
                ("mval", ct.POINTER(ct.c_int)), ("flags", ct.c_int),
                ("mvstart", ct.POINTER(ct.c_ssize_t)),
                ("stride", ct.c_ssize_t)]


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


class SharpLike:
    """Thin wrapper over a libsharp-compatible C interface."""

    def __init__(self, path, prefix, mode=ct.RTLD_LOCAL):
        self.path = path
        self.lib = ct.CDLL(path, mode=mode)
        self.p = prefix
        L = self.lib
        f = self._f
        f("make_geom_info").argtypes = [
            ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
            ct.c_void_p, ct.c_void_p, ct.POINTER(ct.POINTER(GeomInfo))]
        f("make_weighted_healpix_geom_info").argtypes = [
            ct.c_int, ct.c_int, ct.c_void_p, ct.POINTER(ct.POINTER(GeomInfo))]
        for g in ("gauss", "fejer1", "cc", "fejer2", "mw"):
            f("make_%s_geom_info" % g).argtypes = [
                ct.c_int, ct.c_int, ct.c_double, ct.c_int, ct.c_int,
                ct.POINTER(ct.POINTER(GeomInfo))]
        f("destroy_geom_info").argtypes = [ct.POINTER(GeomInfo)]
        f("make_triangular_alm_info").argtypes = [
            ct.c_int, ct.c_int, ct.c_int, ct.POINTER(ct.POINTER(AlmInfo))]
        f("make_rectangular_alm_info").argtypes = [
            ct.c_int, ct.c_int, ct.c_int, ct.POINTER(ct.POINTER(AlmInfo))]
        f("make_general_alm_info").argtypes = [
            ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int,
            ct.POINTER(ct.POINTER(AlmInfo))]
        f("destroy_alm_info").argtypes = [ct.POINTER(AlmInfo)]
        f("execute").argtypes = [
            ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.POINTER(GeomInfo),
            ct.POINTER(AlmInfo), ct.c_int, ct.c_int,
            ct.POINTER(ct.c_double), ct.POINTER(ct.c_ulonglong)]
        f("get_mlim").argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double]
        f("get_mlim").restype = ct.c_int
        del L

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    # --- geometry ---------------------------------------------------------
    def healpix(self, nside, weight=None, stride=1):
        g = ct.POINTER(GeomInfo)()
        w = None if weight is None else _ptr(np.ascontiguousarray(weight, np.float64))
        self._f("make_weighted_healpix_geom_info")(nside, stride, w, ct.byref(g))
        return g

    def iso(self, kind, nrings, nphi, phi0=0.0, stride_lon=1, stride_lat=None):
        g = ct.POINTER(GeomInfo)()
        if stride_lat is None:
            stride_lat = nphi
        self._f("make_%s_geom_info" % kind)(nrings, nphi, phi0, stride_lon,
                                            stride_lat, ct.byref(g))
        return g

    def general_geom(self, nph, ofs, stride, phi0, theta, wgt=None):
        g = ct.POINTER(GeomInfo)()
        nph = np.ascontiguousarray(nph, np.int32)
        ofs = np.ascontiguousarray(ofs, np.int64)
        stride = np.ascontiguousarray(stride, np.int32)
        phi0 = np.ascontiguousarray(phi0, np.float64)
        theta = np.ascontiguousarray(theta, np.float64)
        w = None if wgt is None else _ptr(np.ascontiguousarray(wgt, np.float64))
        self._f("make_geom_info")(len(nph), _ptr(nph), _ptr(ofs), _ptr(stride),
                                  _ptr(phi0), _ptr(theta), w, ct.byref(g))
        return g

    def destroy_geom(self, g):
        self._f("destroy_geom_info")(g)

    # --- a_lm -------------------------------------------------------------
    def triangular(self, lmax, mmax=None, stride=1):
        a = ct.POINTER(AlmInfo)()
        self._f("make_triangular_alm_info")(lmax, lmax if mmax is None else mmax,
                                            stride, ct.byref(a))
        return a

    def rectangular(self, lmax, mmax=None, stride=1):
        a = ct.POINTER(AlmInfo)()
        self._f("make_rectangular_alm_info")(lmax, lmax if mmax is None else mmax,
                                             stride, ct.byref(a))
        return a

    def general_alm(self, lmax, mval, mvstart, stride=1, flags=0):
        a = ct.POINTER(AlmInfo)()
        mval = np.ascontiguousarray(mval, np.int32)
        mvstart = np.ascontiguousarray(mvstart, np.int64)
        self._f("make_general_alm_info")(lmax, len(mval), stride, _ptr(mval),
                                         _ptr(mvstart), flags, ct.byref(a))
        return a

    def destroy_alm(self, a):
        self._f("destroy_alm_info")(a)

    # --- transform --------------------------------------------------------
    def execute(self, jobtype, spin, alms, maps, g, a, ntrans, flags):
        """alms/maps: lists of numpy arrays (complex128/float64 for DP,
        complex64/float32 otherwise), modified in place."""
        ap = (ct.c_void_p * len(alms))(*[a_.ctypes.data for a_ in alms])
        mp = (ct.c_void_p * len(maps))(*[m_.ctypes.data for m_ in maps])
        t = ct.c_double(0.0)
        op = ct.c_ulonglong(0)
        self._f("execute")(jobtype, spin, ap, mp, g, a, ntrans, flags,
                           ct.byref(t), ct.byref(op))
        return t.value, op.value

    def mlim(self, lmax, spin, sth, cth):
        return self._f("get_mlim")(lmax, spin, sth, cth)


def load_oracle():
    """The plain-C restatement; raises if it has not been built."""
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle not built: run `make -C oracle oracle`")
    o = SharpLike(ORACLE_SO, "orc_")
    L = o.lib
    L.orc_ylm_spin0.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                ct.c_void_p]
    L.orc_ylm_spin.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_double,
                               ct.c_double, ct.c_void_p, ct.c_void_p]
    L.orc_get_norm.argtypes = [ct.c_int, ct.c_int, ct.c_void_p]
    L.orc_get_d1norm.argtypes = [ct.c_int, ct.c_void_p]
    L.orc_rdft_forward.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p]
    L.orc_rdft_backward.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p]
    return o


def load_ref():
    """The reference build, or None where it has not been built."""
    if not os.path.exists(REF_SO):
        return None
    return SharpLike(REF_SO, "sharp_")


def ylm_spin0(o, lmax, m, cth, sth):
    out = np.zeros(lmax - m + 1)
    o.lib.orc_ylm_spin0(lmax, m, cth, sth, _ptr(out))
    return out


def ylm_spin(o, lmax, m, spin, cth, sth):
    n = lmax - max(m, spin) + 1
    outp = np.zeros(max(n, 0) + 1)
    outm = np.zeros(max(n, 0) + 1)
    o.lib.orc_ylm_spin(lmax, m, spin, cth, sth, _ptr(outp), _ptr(outm))
    return outp[:n], outm[:n]


def rdft_forward(o, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    o.lib.orc_rdft_forward(len(x), _ptr(x), _ptr(out))
    return out


def rdft_backward(o, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    o.lib.orc_rdft_backward(len(x), _ptr(x), _ptr(out))
    return out


# --- helpers shared by the tests -------------------------------------------
def geom_npix(g):
    gi = g.contents
    n = 0
    for i in range(gi.npairs):
        p = gi.pair[i]
        n += p.r1.nph + (p.r2.nph if p.r2.nph > 0 else 0)
    return n


def geom_arrays(g):
    """Pair table as a dict of numpy arrays (field-by-field comparison)."""
    gi = g.contents
    arr = np.ctypeslib.as_array(
        ct.cast(gi.pair, ct.POINTER(ct.c_char)), (gi.npairs * ct.sizeof(RingPair),))
    rec = np.frombuffer(arr.tobytes(), dtype=np.dtype(
        [(h + f, t) for h in ("r1_", "r2_") for f, t in (
            ("theta", "f8"), ("phi0", "f8"), ("weight", "f8"), ("cth", "f8"),
            ("sth", "f8"), ("ofs", "i8"), ("nph", "i4"), ("stride", "i4"))]))
    return rec, gi.nphmax


def random_alm(lmax, mmax, rng, spin=0, ncomp=1):
    """Triangular (Healpix) layout, m=0 imaginary parts zero, l<spin zero."""
    nalm = ((mmax + 1) * (mmax + 2)) // 2 + (mmax + 1) * (lmax - mmax)
    out = []
    for _ in range(ncomp):
        a = np.zeros(nalm, np.complex128)
        idx = 0
        for m in range(mmax + 1):
            n = lmax - m + 1
            re = rng.uniform(-1, 1, n)
            im = rng.uniform(-1, 1, n) if m > 0 else np.zeros(n)
            blk = re + 1j * im
            ls = np.arange(m, lmax + 1)
            blk[ls < spin] = 0
            a[idx:idx + n] = blk
            idx += n
        out.append(a)
    return out

"""Python mirror of libsharp's interface over the MI355X engine (ctypes).

The native library ``libsharp_hip.so`` (built from ``csrc/``) exports the
libsharp C interface unchanged (libsharp/sharp_lowlevel.h,
sharp_geomhelpers.h, sharp_almhelpers.h) plus the device-resident entry points
of ``include/sharp_hip.h``.  This module binds both:

* ``Geometry`` / ``AlmInfo`` own ``sharp_geom_info`` / ``sharp_alm_info``
  objects built by the library's own constructors;
* ``SharpJob`` mirrors the C++ wrapper ``sharp_cxxjob<T>`` of
  libsharp/sharp_cxx.h:102-149 (alm2map, alm2map_spin, alm2map_der1, map2alm,
  map2alm_spin) on numpy arrays in host memory;
* ``execute`` is ``sharp_execute`` on host arrays and ``execute_device`` its
  device-resident form on torch CUDA tensors (no host round trip).

If the library is missing this module raises at import: there is no CPU
fallback.
"""
import ctypes as ct
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SHARP_HIP_LIBRARY names another build of the library (e.g. one made with
# different build-time options); default: the in-tree build
LIB_PATH = os.environ.get("SHARP_HIP_LIBRARY") or os.path.join(HERE, "libsharp_hip.so")

# job types and flags (libsharp/sharp_lowlevel.h:162-191, include/sharp_hip.h)
MAP2ALM, ALM2MAP, Yt, WY, ALM2MAP_DERIV1 = 0, 1, 2, 3, 4
DP, ADD, REAL_HARMONICS, NO_FFT = 1 << 4, 1 << 5, 1 << 6, 1 << 7
MIXED_FP32, MIXED_BF16X3 = 1 << 24, 1 << 25
PACKED = 1


class RingInfo(ct.Structure):
    _fields_ = [("theta", ct.c_double), ("phi0", ct.c_double),
                ("weight", ct.c_double), ("cth", ct.c_double),
                ("sth", ct.c_double), ("ofs", ct.c_ssize_t),
                ("nph", ct.c_int), ("stride", ct.c_int)]


class RingPair(ct.Structure):
    _fields_ = [("r1", RingInfo), ("r2", RingInfo)]


class GeomInfoC(ct.Structure):
    _fields_ = [("pair", ct.POINTER(RingPair)), ("npairs", ct.c_int),
                ("nphmax", ct.c_int)]


class AlmInfoC(ct.Structure):
    _fields_ = [("lmax", ct.c_int), ("nm", ct.c_int),
                ("mval", ct.POINTER(ct.c_int)), ("flags", ct.c_int),
                ("mvstart", ct.POINTER(ct.c_ssize_t)),
                ("stride", ct.c_ssize_t)]


def _load():
    # One HIP runtime per process: PyTorch's ROCm wheels carry their own
    # libamdhip64 (with the same soname); when torch is present it is loaded
    # first, so this library binds to that runtime instead of a second copy.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libsharp_hip.so is not built (run __graft_entry__.build() or "
            "make -C planck-levels_amd/csrc); there is no CPU fallback")
    lib = ct.CDLL(LIB_PATH, mode=ct.RTLD_LOCAL)
    G = ct.POINTER(GeomInfoC)
    A = ct.POINTER(AlmInfoC)
    lib.sharp_make_geom_info.argtypes = [
        ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
        ct.c_void_p, ct.c_void_p, ct.POINTER(G)]
    lib.sharp_make_weighted_healpix_geom_info.argtypes = [
        ct.c_int, ct.c_int, ct.c_void_p, ct.POINTER(G)]
    for g in ("gauss", "fejer1", "cc", "fejer2", "mw"):
        getattr(lib, "sharp_make_%s_geom_info" % g).argtypes = [
            ct.c_int, ct.c_int, ct.c_double, ct.c_int, ct.c_int, ct.POINTER(G)]
    lib.sharp_destroy_geom_info.argtypes = [G]
    lib.sharp_make_triangular_alm_info.argtypes = [
        ct.c_int, ct.c_int, ct.c_int, ct.POINTER(A)]
    lib.sharp_make_rectangular_alm_info.argtypes = [
        ct.c_int, ct.c_int, ct.c_int, ct.POINTER(A)]
    lib.sharp_make_general_alm_info.argtypes = [
        ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int,
        ct.POINTER(A)]
    lib.sharp_destroy_alm_info.argtypes = [A]
    lib.sharp_alm_index.argtypes = [A, ct.c_int, ct.c_int]
    lib.sharp_alm_index.restype = ct.c_ssize_t
    lib.sharp_execute.argtypes = [
        ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, G, A, ct.c_int, ct.c_int,
        ct.POINTER(ct.c_double), ct.POINTER(ct.c_ulonglong)]
    lib.sharp_hip_execute_device.argtypes = [
        ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, G, A, ct.c_int, ct.c_int,
        ct.c_void_p, ct.POINTER(ct.c_ulonglong)]
    lib.sharp_hip_execute_device_noalloc.argtypes = [
        ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, G, A, ct.c_int, ct.c_int,
        ct.c_void_p, ct.POINTER(ct.c_ulonglong)]
    for f in ("sharp_hip_alm2map_adjoint_device",
              "sharp_hip_map2alm_adjoint_device"):
        getattr(lib, f).argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, G, A,
                                    ct.c_int, ct.c_int, ct.c_void_p]
    lib.sharp_hip_execute_scaled_device.argtypes = [
        ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, G, A, ct.c_int, ct.c_int,
        ct.POINTER(ct.c_double), ct.c_void_p]
    lib.sharp_hip_map2alm_iter_device.argtypes = [
        ct.c_int, ct.c_void_p, ct.c_void_p, G, A, ct.c_int, ct.c_int, ct.c_int,
        ct.c_void_p]
    lib.sharp_hip_batch_create.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.c_int]
    lib.sharp_hip_batch_create.restype = ct.c_void_p
    lib.sharp_hip_batch_execute.argtypes = [ct.c_void_p, ct.c_void_p]
    lib.sharp_hip_batch_config.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int),
                                           ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
    lib.sharp_hip_batch_tiles.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
    lib.sharp_hip_batch_destroy.argtypes = [ct.c_void_p]
    lib.sharp_hip_prepare.argtypes = [G, A, ct.c_int]
    lib.sharp_hip_synchronize.argtypes = [ct.c_void_p]
    lib.sharp_hip_device_count.restype = ct.c_int
    lib.sharp_hip_profile.argtypes = [ct.c_int]
    lib.sharp_hip_profile_stats.argtypes = [
        ct.c_int, ct.c_char_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    lib.sharp_hip_profile_stats.restype = ct.c_int
    lib.sharp_get_mlim.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double]
    lib.sharp_get_mlim.restype = ct.c_int
    return lib


LIB = _load()


def _np_ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


class Geometry:
    """Owns a sharp_geom_info built by the library (sharp_geomhelpers.h)."""

    def __init__(self, ptr):
        self.ptr = ptr

    @classmethod
    def healpix(cls, nside, weight=None, stride=1):
        g = ct.POINTER(GeomInfoC)()
        w = None if weight is None else _np_ptr(np.ascontiguousarray(weight, np.float64))
        LIB.sharp_make_weighted_healpix_geom_info(nside, stride, w, ct.byref(g))
        return cls(g)

    @classmethod
    def iso(cls, kind, nrings, nphi, phi0=0.0, stride_lon=1, stride_lat=None):
        """kind: 'gauss', 'fejer1' (= ECP), 'cc', 'fejer2', 'mw'."""
        g = ct.POINTER(GeomInfoC)()
        getattr(LIB, "sharp_make_%s_geom_info" % kind)(
            nrings, nphi, phi0, stride_lon,
            nphi if stride_lat is None else stride_lat, ct.byref(g))
        return cls(g)

    @classmethod
    def general(cls, nph, ofs, stride, phi0, theta, wgt=None):
        g = ct.POINTER(GeomInfoC)()
        arrs = [np.ascontiguousarray(nph, np.int32), np.ascontiguousarray(ofs, np.int64),
                np.ascontiguousarray(stride, np.int32), np.ascontiguousarray(phi0, np.float64),
                np.ascontiguousarray(theta, np.float64)]
        w = None if wgt is None else np.ascontiguousarray(wgt, np.float64)
        LIB.sharp_make_geom_info(len(arrs[0]), *[_np_ptr(a) for a in arrs],
                                 None if w is None else _np_ptr(w), ct.byref(g))
        g._keep = (arrs, w)
        return cls(g)

    @property
    def npix(self):
        gi = self.ptr.contents
        n = 0
        for i in range(gi.npairs):
            p = gi.pair[i]
            n += p.r1.nph + (p.r2.nph if p.r2.nph > 0 else 0)
        return n

    def __del__(self):
        if getattr(self, "ptr", None):
            LIB.sharp_destroy_geom_info(self.ptr)
            self.ptr = None


class AlmInfo:
    """Owns a sharp_alm_info (sharp_almhelpers.h / sharp_lowlevel.h)."""

    def __init__(self, ptr):
        self.ptr = ptr

    @classmethod
    def triangular(cls, lmax, mmax=None, stride=1):
        a = ct.POINTER(AlmInfoC)()
        LIB.sharp_make_triangular_alm_info(lmax, lmax if mmax is None else mmax,
                                           stride, ct.byref(a))
        return cls(a)

    @classmethod
    def rectangular(cls, lmax, mmax=None, stride=1):
        a = ct.POINTER(AlmInfoC)()
        LIB.sharp_make_rectangular_alm_info(lmax, lmax if mmax is None else mmax,
                                            stride, ct.byref(a))
        return cls(a)

    @classmethod
    def general(cls, lmax, mval, mvstart, stride=1, flags=0):
        a = ct.POINTER(AlmInfoC)()
        mval = np.ascontiguousarray(mval, np.int32)
        mvstart = np.ascontiguousarray(mvstart, np.int64)
        LIB.sharp_make_general_alm_info(lmax, len(mval), stride, _np_ptr(mval),
                                        _np_ptr(mvstart), flags, ct.byref(a))
        return cls(a)

    @property
    def lmax(self):
        return self.ptr.contents.lmax

    @property
    def mmax(self):
        return self.ptr.contents.nm - 1

    def index(self, l, mi):
        return LIB.sharp_alm_index(self.ptr, l, mi)

    def __del__(self):
        if getattr(self, "ptr", None):
            LIB.sharp_destroy_alm_info(self.ptr)
            self.ptr = None


def nalm_triangular(lmax, mmax=None):
    mmax = lmax if mmax is None else mmax
    return ((mmax + 1) * (mmax + 2)) // 2 + (mmax + 1) * (lmax - mmax)


def execute(jobtype, spin, alms, maps, geom, ainfo, ntrans, flags):
    """sharp_execute on numpy arrays (modified in place).

    Returns (wall time in s, operation count of libsharp's flop model)."""
    ap = (ct.c_void_p * len(alms))(*[a.ctypes.data for a in alms])
    mp = (ct.c_void_p * len(maps))(*[m.ctypes.data for m in maps])
    t = ct.c_double(0.0)
    op = ct.c_ulonglong(0)
    LIB.sharp_execute(jobtype, spin, ap, mp, geom.ptr, ainfo.ptr, ntrans, flags,
                      ct.byref(t), ct.byref(op))
    return t.value, op.value


def execute_device(jobtype, spin, alms, maps, geom, ainfo, ntrans, flags,
                   stream=None):
    """Device-resident sharp_execute on torch CUDA tensors.

    alms: complex128/complex64 tensors (or real views thereof) in the layout
    described by ainfo; maps: float64/float32 tensors.  Work is enqueued on
    `stream` (default: torch's current stream); returns the operation count."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    ap = (ct.c_void_p * len(alms))(*[a.data_ptr() for a in alms])
    mp = (ct.c_void_p * len(maps))(*[m.data_ptr() for m in maps])
    op = ct.c_ulonglong(0)
    LIB.sharp_hip_execute_device(jobtype, spin, ap, mp, geom.ptr, ainfo.ptr,
                                 ntrans, flags, ct.c_void_p(stream.cuda_stream),
                                 ct.byref(op))
    return op.value


def prepare(geom, ainfo, spin):
    """Builds the device tables of (geometry, a_lm layout, spin) ahead of the
    first transform (sharp_hip_prepare)."""
    LIB.sharp_hip_prepare(geom.ptr, ainfo.ptr, spin)


def execute_device_noalloc(jobtype, spin, alms, maps, geom, ainfo, ntrans, flags,
                           stream=None):
    """execute_device through sharp_hip_execute_device_noalloc: no device
    allocation; needs prepare(geom, ainfo, spin) first and a job whose work
    memory fits the library's fixed scratch area."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    ap = (ct.c_void_p * len(alms))(*[a.data_ptr() for a in alms])
    mp = (ct.c_void_p * len(maps))(*[m.data_ptr() for m in maps])
    op = ct.c_ulonglong(0)
    LIB.sharp_hip_execute_device_noalloc(jobtype, spin, ap, mp, geom.ptr, ainfo.ptr,
                                         ntrans, flags, ct.c_void_p(stream.cuda_stream),
                                         ct.byref(op))
    return op.value


def _dev_call(fn, spin, alms, maps, geom, ainfo, ntrans, flags, stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    ap = (ct.c_void_p * len(alms))(*[a.data_ptr() for a in alms])
    mp = (ct.c_void_p * len(maps))(*[m.data_ptr() for m in maps])
    fn(spin, ap, mp, geom.ptr, ainfo.ptr, ntrans, flags, ct.c_void_p(stream.cuda_stream))


def alm2map_adjoint_device(spin, alms, maps, geom, ainfo, ntrans, flags, stream=None):
    """Adjoint of synthesis on device tensors: alms = Y^T maps (map2alm without
    quadrature weights, libsharp's SHARP_Yt).  alms are overwritten (or
    accumulated with ADD)."""
    _dev_call(LIB.sharp_hip_alm2map_adjoint_device, spin, alms, maps, geom, ainfo,
              ntrans, flags, stream)


def map2alm_adjoint_device(spin, alms, maps, geom, ainfo, ntrans, flags, stream=None):
    """Adjoint of analysis on device tensors: maps = W Y alms (alm2map with
    quadrature weights, libsharp's SHARP_WY)."""
    _dev_call(LIB.sharp_hip_map2alm_adjoint_device, spin, alms, maps, geom, ainfo,
              ntrans, flags, stream)


def execute_scaled_device(jobtype, spin, alms, maps, geom, ainfo, ntrans, flags, lfac,
                          stream=None):
    """execute_device with the per-l factor lfac (lmax+1 values) on the a_lm
    side: synthesis maps = Y (lfac_l a_lm) (the caller's a_lm unchanged),
    analysis a_lm (+)= lfac_l Y^T W maps."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    f = np.ascontiguousarray(lfac, dtype=np.float64)
    if f.size < ainfo.lmax + 1:
        raise ValueError("lfac needs lmax+1 values")
    ap = (ct.c_void_p * len(alms))(*[a.data_ptr() for a in alms])
    mp = (ct.c_void_p * len(maps))(*[m.data_ptr() for m in maps])
    LIB.sharp_hip_execute_scaled_device(
        jobtype, spin, ap, mp, geom.ptr, ainfo.ptr, ntrans, flags,
        f.ctypes.data_as(ct.POINTER(ct.c_double)), ct.c_void_p(stream.cuda_stream))


def map2alm_iter_device(spin, alms, maps, geom, ainfo, ntrans, flags, num_iter,
                        stream=None):
    """Jacobi-iterated analysis on device tensors (Healpix_cxx map2alm_iter):
    alms = Y^T W maps, then num_iter times alms += Y^T W (maps - Y alms)."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    ap = (ct.c_void_p * len(alms))(*[a.data_ptr() for a in alms])
    mp = (ct.c_void_p * len(maps))(*[m.data_ptr() for m in maps])
    LIB.sharp_hip_map2alm_iter_device(spin, ap, mp, geom.ptr, ainfo.ptr, ntrans, flags,
                                      num_iter, ct.c_void_p(stream.cuda_stream))


FWHM2SIGMA = 1.0 / math.sqrt(8.0 * math.log(2.0))


def gauss_beam(lmax, fwhm, pol=False):
    """Per-l factors of Healpix's smoothWithGauss (Healpix_cxx/
    alm_powspec_tools.cc:250-283): exp(-l(l+1) sigma^2 / 2) with sigma =
    |fwhm| * fwhm2sigma (radians), the inverse for fwhm < 0, and the extra
    exp(2 sigma^2) (or its inverse) of the polarisation a_lm when pol."""
    fct = 1 if fwhm >= 0 else -1
    sigma = fwhm * FWHM2SIGMA
    l = np.arange(lmax + 1, dtype=np.float64)
    gb = np.exp(-0.5 * fct * l * (l + 1) * sigma * sigma)
    if pol:
        gb = gb * math.exp(2 * fct * sigma * sigma)
    return gb


def smoothing_device(maps, geom, ainfo, spin, fwhm, num_iter, alms, flags=None,
                     stream=None):
    """smoothing_cxx on device tensors (Healpix_cxx/smoothing_cxx_module.cc:
    79-81, 108-111) for one spin: alms = map2alm_iter(maps), then
    maps = alm2map(beam_l * alms) with the Gaussian beam of `fwhm` (radians)
    applied while the a_lm are gathered.  `alms` receive the unsmoothed a_lm.
    ntrans = len(maps) / (1 for spin 0, 2 otherwise)."""
    if flags is None:
        flags = DP
    nc = 1 if spin == 0 else 2
    ntrans = len(maps) // nc
    map2alm_iter_device(spin, alms, maps, geom, ainfo, ntrans, flags, num_iter, stream)
    execute_scaled_device(ALM2MAP, spin, alms, maps, geom, ainfo, ntrans, flags,
                          gauss_beam(ainfo.lmax, fwhm, pol=spin != 0), stream)


class _JobC(ct.Structure):
    _fields_ = [("type", ct.c_int), ("spin", ct.c_int), ("alm", ct.c_void_p),
                ("map", ct.c_void_p), ("geom_info", ct.POINTER(GeomInfoC)),
                ("alm_info", ct.POINTER(AlmInfoC)), ("ntrans", ct.c_int)]


class Batch:
    """A batch of device transforms of uneven size (include/sharp_hip.h,
    sharp_hip_batch_*): jobs is a list of (jobtype, spin, alms, maps, geom,
    ainfo, ntrans) with torch CUDA tensors as in execute_device.  The tensors
    must stay alive (and in place) for the life of the batch; with use_graph
    each execute() replays the HIP graph recorded at creation."""

    def __init__(self, jobs, flags, use_graph=True):
        self._keep = []
        arr = (_JobC * len(jobs))()
        for k, (jt, spin, alms, maps, geom, ainfo, ntrans) in enumerate(jobs):
            ap = (ct.c_void_p * len(alms))(*[a.data_ptr() for a in alms])
            mp = (ct.c_void_p * len(maps))(*[m.data_ptr() for m in maps])
            self._keep += [ap, mp, alms, maps, geom, ainfo]
            arr[k] = _JobC(jt, spin, ct.cast(ap, ct.c_void_p), ct.cast(mp, ct.c_void_p),
                           geom.ptr, ainfo.ptr, ntrans)
        self._arr = arr
        self.ptr = LIB.sharp_hip_batch_create(len(jobs), ct.cast(arr, ct.c_void_p), flags,
                                              1 if use_graph else 0)

    def execute(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        LIB.sharp_hip_batch_execute(self.ptr, ct.c_void_p(stream.cuda_stream))

    def config(self):
        t, w, g = ct.c_int(), ct.c_int(), ct.c_int()
        LIB.sharp_hip_batch_config(self.ptr, ct.byref(t), ct.byref(w), ct.byref(g))
        mb, rw = ct.c_int(), ct.c_int()
        LIB.sharp_hip_batch_tiles(self.ptr, ct.byref(mb), ct.byref(rw))
        return {"transforms_per_tile": t.value, "workgroups_per_cu": w.value,
                "graph": bool(g.value), "m_block": mb.value, "ring_waves": rw.value}

    def close(self):
        if getattr(self, "ptr", None) and LIB is not None:
            LIB.sharp_hip_batch_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SharpJob:
    """Mirror of sharp_cxxjob<T> (libsharp/sharp_cxx.h:39-149) on numpy arrays;
    dtype float64 (SHARP_DP) or float32."""

    def __init__(self, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.geom = None
        self.ainfo = None

    def _flags(self, add):
        return (DP if self.dtype == np.float64 else 0) | (ADD if add else 0)

    def set_Healpix_geometry(self, nside):
        self.geom = Geometry.healpix(nside)

    def set_weighted_Healpix_geometry(self, nside, weight):
        self.geom = Geometry.healpix(nside, weight)

    def set_Gauss_geometry(self, nrings, nphi):
        self.geom = Geometry.iso("gauss", nrings, nphi)

    def set_ECP_geometry(self, nrings, nphi):
        self.geom = Geometry.iso("fejer1", nrings, nphi)

    def set_general_geometry(self, nph, ofs, stride, phi0, theta, wgt=None):
        self.geom = Geometry.general(nph, ofs, stride, phi0, theta, wgt)

    def set_triangular_alm_info(self, lmax, mmax):
        self.ainfo = AlmInfo.triangular(lmax, mmax)

    def alm2map(self, alm, map_, add=False):
        execute(ALM2MAP, 0, [alm], [map_], self.geom, self.ainfo, 1, self._flags(add))

    def alm2map_spin(self, alm1, alm2, map1, map2, spin, add=False):
        execute(ALM2MAP, spin, [alm1, alm2], [map1, map2], self.geom, self.ainfo,
                1, self._flags(add))

    def alm2map_der1(self, alm, map1, map2, add=False):
        execute(ALM2MAP_DERIV1, 1, [alm], [map1, map2], self.geom, self.ainfo, 1,
                self._flags(add))

    def map2alm(self, map_, alm, add=False):
        execute(MAP2ALM, 0, [alm], [map_], self.geom, self.ainfo, 1, self._flags(add))

    def map2alm_spin(self, map1, map2, alm1, alm2, spin, add=False):
        execute(MAP2ALM, spin, [alm1, alm2], [map1, map2], self.geom, self.ainfo,
                1, self._flags(add))


def device_count():
    return LIB.sharp_hip_device_count()


def release():
    """Frees the cached tables and work buffers of the current device (live
    batches keep what their jobs use)."""
    LIB.sharp_hip_release()


def synchronize(stream=None):
    LIB.sharp_hip_synchronize(None if stream is None else ct.c_void_p(stream.cuda_stream))


def profile(enable):
    """Start (True, clears counters) or stop per-stage event timing."""
    LIB.sharp_hip_profile(1 if enable else 0)


def profile_stats(maxn=64):
    """Per-stage totals recorded since profile(True): list of dicts with
    name, ms, launches, flops (libsharp flop model), bytes (compulsory HBM)."""
    names = ct.create_string_buffer(32 * maxn)
    ms = np.zeros(maxn)
    launches = np.zeros(maxn, np.int64)
    flops = np.zeros(maxn)
    byts = np.zeros(maxn)
    n = LIB.sharp_hip_profile_stats(maxn, names, _np_ptr(ms), _np_ptr(launches),
                                    _np_ptr(flops), _np_ptr(byts))
    out = []
    for i in range(min(n, maxn)):
        nm = names.raw[32 * i:32 * (i + 1)].split(b"\0", 1)[0].decode()
        out.append(dict(name=nm, ms=float(ms[i]), launches=int(launches[i]),
                        flops=float(flops[i]), bytes=float(byts[i])))
    return out

"""torch.autograd over the spherical harmonic transforms of the HIP engine.

``alm2map`` (synthesis, Y) takes its gradient from the hand-written adjoint
Y^T (libsharp's SHARP_Yt: analysis without quadrature weights) and
``map2alm`` (analysis, Y^T W) from W Y (SHARP_WY), so gradients of map-domain
losses with respect to a_lm need no dense matrices.

a_lm enter and leave as real tensors of shape (nalm, 2) -- (Re, Im) of the
complex128 array in the layout of `ainfo`, i.e. ``torch.view_as_real`` of it
-- so that every real degree of freedom has an ordinary gradient.  Real maps
hold only m >= 0 coefficients: a map is sum_m c_m Re(a_lm Y_lm) with c_0 = 1
(only Re a_l0 enters) and c_m = 2 for m > 0 (libsharp/sharp_core_inc2.c:
32-117, with the m=0 imaginary part ignored, libsharp/sharp.c:481-606).
Hence for a loss L(map):

    dL/dRe a_lm = c_m Re[Y^T g]_lm,  dL/dIm a_lm = c_m Im[Y^T g]_lm  (m > 0),
    dL/dIm a_l0 = 0,                                 g = dL/dmap,

and for L(a) with a = map2alm(x): dL/dx = W Y g' with g'_lm = (gRe + i gIm)/c_m.

Spin s > 0 works the same way on (E, B) <-> (Q, U) pairs.  The transforms
themselves are supplied by an engine object; the default engine is the HIP
library on the current device (there is no CPU engine in the product).
"""
import numpy as np
import torch

import sharp_hip as S


def m_weights(ainfo, nalm, device):
    """c_m for every complex a_lm entry of the layout (1 for m=0, 2 for m>0);
    entries no (l, m) maps to keep 0."""
    ai = ainfo.ptr.contents
    c = np.zeros(nalm, dtype=np.float64)
    for mi in range(ai.nm):
        m = ai.mval[mi]
        idx = ai.mvstart[mi] + np.arange(m, ai.lmax + 1) * ai.stride
        c[idx] = 1.0 if m == 0 else 2.0
    return torch.from_numpy(c).to(device)


class DeviceEngine:
    """The HIP library on torch CUDA tensors (FP64, or the mixed FP32 form)."""

    def __init__(self, geom, ainfo, spin, mixed=False):
        self.geom, self.ainfo, self.spin = geom, ainfo, spin
        self.flags = S.DP | (S.MIXED_FP32 if mixed else 0)

    def synth(self, alms, maps):
        S.execute_device(S.ALM2MAP, self.spin, alms, maps, self.geom, self.ainfo,
                         1, self.flags)

    def synth_adjoint(self, maps, alms):
        S.alm2map_adjoint_device(self.spin, alms, maps, self.geom, self.ainfo, 1,
                                 self.flags)

    def anal(self, maps, alms):
        S.execute_device(S.MAP2ALM, self.spin, alms, maps, self.geom, self.ainfo,
                         1, self.flags)

    def anal_adjoint(self, alms, maps):
        S.map2alm_adjoint_device(self.spin, alms, maps, self.geom, self.ainfo, 1,
                                 self.flags)


def _ncomp(spin):
    return 1 if spin == 0 else 2


class _Alm2Map(torch.autograd.Function):
    @staticmethod
    def forward(ctx, engine, npix, cm, *alms):
        ctx.engine, ctx.cm = engine, cm
        calms = [torch.view_as_complex(a.contiguous()) for a in alms]
        maps = [torch.zeros(npix, dtype=torch.float64, device=alms[0].device)
                for _ in calms]
        engine.synth(calms, maps)
        return tuple(maps)

    @staticmethod
    def backward(ctx, *gmaps):
        e, cm = ctx.engine, ctx.cm
        g = [x.contiguous() if x is not None else None for x in gmaps]
        ref = next(x for x in g if x is not None)
        g = [x if x is not None else torch.zeros_like(ref) for x in g]
        out = [torch.zeros(cm.shape[0], dtype=torch.complex128, device=ref.device)
               for _ in g]
        e.synth_adjoint(g, out)
        grads = []
        for o in out:
            r = torch.view_as_real(o) * cm[:, None]
            r[:, 1] = torch.where(cm == 1.0, torch.zeros_like(r[:, 1]), r[:, 1])
            grads.append(r)
        return (None, None, None) + tuple(grads)


class _Map2Alm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, engine, cm, *maps):
        ctx.engine, ctx.cm, ctx.npix = engine, cm, maps[0].shape[0]
        out = [torch.zeros(cm.shape[0], dtype=torch.complex128, device=maps[0].device)
               for _ in maps]
        engine.anal([m.contiguous() for m in maps], out)
        return tuple(torch.view_as_real(o) for o in out)

    @staticmethod
    def backward(ctx, *galms):
        e, cm = ctx.engine, ctx.cm
        inv = torch.where(cm > 0, 1.0 / torch.clamp(cm, min=1.0), torch.zeros_like(cm))
        g = []
        for x in galms:
            if x is None:
                x = torch.zeros(cm.shape[0], 2, dtype=torch.float64, device=cm.device)
            g.append(torch.view_as_complex((x * inv[:, None]).contiguous()))
        maps = [torch.zeros(ctx.npix, dtype=torch.float64, device=cm.device) for _ in g]
        e.anal_adjoint(g, maps)
        return (None, None) + tuple(maps)


class SHT:
    """Differentiable transforms for one geometry / a_lm layout / spin.

    alm2map(*alms) -> maps and map2alm(*maps) -> alms, alms as (nalm, 2)
    real tensors (one for spin 0, (E, B) for spin s), maps as (npix,) float64
    tensors."""

    def __init__(self, geom, ainfo, spin=0, engine=None, device=None, mixed=False):
        self.geom, self.ainfo, self.spin = geom, ainfo, spin
        self.engine = engine if engine is not None else DeviceEngine(geom, ainfo, spin, mixed)
        self.device = torch.device("cuda") if device is None else torch.device(device)
        ai = ainfo.ptr.contents
        self.nalm = max(int(ai.mvstart[mi]) + ai.lmax * int(ai.stride) + 1
                        for mi in range(ai.nm))
        self.npix = geom.npix
        self.cm = m_weights(ainfo, self.nalm, self.device)

    def alm2map(self, *alms):
        assert len(alms) == _ncomp(self.spin)
        return _Alm2Map.apply(self.engine, self.npix, self.cm, *alms)

    def map2alm(self, *maps):
        assert len(maps) == _ncomp(self.spin)
        return _Map2Alm.apply(self.engine, self.cm, *maps)

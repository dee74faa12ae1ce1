"""sharp_execute on host arrays at sizes where the call runs as two groups of
transforms (more than 256 MiB staged through the pinned ring): upload of the
second group and download of the first overlap the device work of the other.
Checked against the same library's device-array entry point (which the parity
tests pin to the C restatement and the reference build) on the same seeded
inputs, with an odd transform count so the groups differ in size, in both
directions, for the FP64 and the BF16-split forms, and against libsharp's
operation count of the whole call."""
import numpy as np
import pytest
import torch

import sharp_hip as S
from sht_inputs import make_alms, spin0_alms

pytestmark = pytest.mark.gpu

NSIDE, LMAX = 512, 1024


def _rel(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


@pytest.mark.parametrize("form", [0, S.MIXED_BF16X3], ids=["fp64", "bf16x3"])
@pytest.mark.parametrize("spin,ntrans", [(2, 7), (0, 13)])
def test_host_call_in_two_groups(gpu, form, spin, ntrans):
    g, a = S.Geometry.healpix(NSIDE), S.AlmInfo.triangular(LMAX)
    npix = 12 * NSIDE * NSIDE
    nm = 2 if spin else 1
    flags = S.DP | form
    # 7 spin-2 transforms: 14 maps of 25 MB; 13 spin-0 ones: 13 maps
    assert ntrans * nm * npix * 8 > 256 << 20
    if spin:
        _, E, B = make_alms(torch, LMAX, ntrans, seed=11, dev=gpu)
        alms_d = [x for t in range(ntrans) for x in (E[t], B[t])]
    else:
        gen = torch.Generator(device=gpu)
        gen.manual_seed(12)
        alms_d = spin0_alms(torch, LMAX, ntrans, gen, gpu)
    alms_h = [x.cpu().numpy() for x in alms_d]

    # synthesis: host arrays vs device arrays
    maps_d = [torch.zeros(npix, dtype=torch.float64, device=gpu) for _ in range(ntrans * nm)]
    op_d = S.execute_device(S.ALM2MAP, spin, alms_d, maps_d, g, a, ntrans, flags)
    maps_h = [np.zeros(npix) for _ in range(ntrans * nm)]
    _, op_h = S.execute(S.ALM2MAP, spin, alms_h, maps_h, g, a, ntrans, flags)
    torch.cuda.synchronize()
    assert op_h == op_d
    err = max(_rel(h, d.cpu().numpy()) for h, d in zip(maps_h, maps_d))
    assert err < 1e-13, err

    # analysis of those maps: host arrays vs device arrays
    out_d = [torch.zeros_like(alms_d[0]) for _ in range(ntrans * nm)]
    op_d = S.execute_device(S.MAP2ALM, spin, out_d, maps_d, g, a, ntrans, flags)
    out_h = [np.zeros_like(alms_h[0]) for _ in range(ntrans * nm)]
    _, op_h = S.execute(S.MAP2ALM, spin, out_h, maps_h, g, a, ntrans, flags)
    torch.cuda.synchronize()
    assert op_h == op_d
    err = max(_rel(h, d.cpu().numpy()) for h, d in zip(out_h, out_d))
    assert err < 1e-13, err

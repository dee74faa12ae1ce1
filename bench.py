"""Benchmark: alm2map + map2alm of TQU maps on one MI355X.

Workload (default): HEALPix Nside=2048, lmax=mmax=4096, a batch of 8
independent TQU realisations.  One step = one round trip of the batch:
  alm2map spin 0 (T) and spin 2 (E,B -> Q,U), then map2alm of both,
all through the device-resident entry point of the drop-in library with the
inputs resident in HBM.  value = TQU round trips per second.

Inputs are synthetic: seeded Gaussian a_lm with C_l^TT = 1/(l(l+1)+1) times a
5-arcmin Gaussian beam window b_l^2, C^EE = 0.1 C^TT, C^BB = 0.01 C^TT and a TE
correlation coefficient of 0.5.

The timed region records HIP events around every stage (on the stream the
kernels run on); the slowest stage gives the roofline object.  The CPU
baseline runs the reference libsharp (oracle/_ref, OpenMP) on one realisation
of the same workload in the same run, when that build is present.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fp64]
"""
import argparse
import ctypes as ct
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "planck-levels_amd"))

METRIC = ("alm2map+map2alm/s, Nside 2048 lmax 4096 TQU; mixed vs FP64 at 1e-6 "
          "rel err")
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "hbm": 8.0e12}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spectra(lmax, fwhm_arcmin=5.0):
    l = np.arange(lmax + 1, dtype=np.float64)
    sigma = fwhm_arcmin / 60.0 * math.pi / 180.0 / math.sqrt(8.0 * math.log(2.0))
    bl2 = np.exp(-l * (l + 1) * sigma * sigma)
    tt = bl2 / (l * (l + 1) + 1)
    ee = 0.1 * tt
    bb = 0.01 * tt
    te = 0.5 * np.sqrt(tt * ee)
    return tt, ee, bb, te


def make_alms(torch, lmax, batch, seed, dev):
    """Triangular (Healpix) layout, complex128, on the GPU."""
    tt, ee, bb, te = spectra(lmax)
    m_of = np.concatenate([np.full(lmax + 1 - m, m) for m in range(lmax + 1)])
    l_of = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    l_t = torch.from_numpy(l_of).to(dev)
    m0 = torch.from_numpy(m_of == 0).to(dev)
    ctt = torch.from_numpy(tt).to(dev)[l_t]
    cee = torch.from_numpy(ee).to(dev)[l_t]
    cbb = torch.from_numpy(bb).to(dev)[l_t]
    cte = torch.from_numpy(te).to(dev)[l_t]
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    n = len(l_of)

    def cgauss():
        re = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        im = torch.randn(n, generator=g, device=dev, dtype=torch.float64)
        # m>0: Re, Im ~ N(0, 1/2); m=0: real N(0, 1)
        s = torch.where(m0, torch.ones_like(re), torch.full_like(re, math.sqrt(0.5)))
        im = torch.where(m0, torch.zeros_like(im), im)
        return torch.complex(re * s, im * s)

    T, E, B = [], [], []
    for _ in range(batch):
        g1, g2, g3 = cgauss(), cgauss(), cgauss()
        at = torch.sqrt(ctt) * g1
        ae = (cte / torch.sqrt(ctt)) * g1 + torch.sqrt(torch.clamp(cee - cte * cte / ctt, min=0)) * g2
        ab = torch.sqrt(cbb) * g3
        T.append(at.contiguous())
        E.append(ae.contiguous())
        B.append(ab.contiguous())
    return T, E, B


def cpu_baseline(nside, lmax, T0, E0, B0):
    """Reference libsharp (OpenMP) on one TQU realisation: alm2map + map2alm.
    Returns (dict, maps, alms) or (None, None, None) without the build."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.load_ref()
    if ref is None:
        return None, None, None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    g = ref.healpix(nside)
    a = ref.triangular(lmax)
    npix = 12 * nside * nside
    alm_t = [T0.copy()]
    alm_eb = [E0.copy(), B0.copy()]
    map_t = [np.zeros(npix)]
    map_qu = [np.zeros(npix), np.zeros(npix)]
    out_t = [np.zeros_like(T0)]
    out_eb = [np.zeros_like(E0), np.zeros_like(B0)]
    # warm-up (fills libsharp's nv timing oracle) on a small job
    gs = ref.healpix(16)
    as_ = ref.triangular(32)
    for sp, na in ((0, 1), (2, 2)):
        ref.execute(O.ALM2MAP, sp, [np.zeros(561, np.complex128)] * na,
                    [np.zeros(3072)] * na, gs, as_, 1, O.DP)
        ref.execute(O.MAP2ALM, sp, [np.zeros(561, np.complex128) for _ in range(na)],
                    [np.zeros(3072)] * na, gs, as_, 1, O.DP)
    t0 = time.perf_counter()
    ref.execute(O.ALM2MAP, 0, alm_t, map_t, g, a, 1, O.DP)
    ref.execute(O.ALM2MAP, 2, alm_eb, map_qu, g, a, 1, O.DP)
    ref.execute(O.MAP2ALM, 0, out_t, map_t, g, a, 1, O.DP)
    ref.execute(O.MAP2ALM, 2, out_eb, map_qu, g, a, 1, O.DP)
    dt = time.perf_counter() - t0
    info = {"value": 1.0 / dt, "unit": "TQU round trips/s", "cores": threads,
            "kind": "reference",
            "sample": "one TQU realisation (1 of the batch of 8): libsharp "
                      "alm2map spin 0+2 then map2alm spin 0+2 at Nside %d lmax %d, "
                      "%.1f s" % (nside, lmax, dt)}
    return info, (map_t[0], map_qu[0], map_qu[1]), (out_t[0], out_eb[0], out_eb[1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--form", default="fp64", choices=["fp64"])
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (calibrates HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the multi-GPU form of this path "
              "is not built (see DESIGN.md); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import sharp_hip as S

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nside = args.nside
    lmax = args.lmax if args.lmax is not None else 2 * nside
    batch = args.batch
    npix = 12 * nside * nside
    geom = S.Geometry.healpix(nside)
    ainfo = S.AlmInfo.triangular(lmax)
    flags = S.DP

    log("generating inputs: nside %d lmax %d batch %d" % (nside, lmax, batch))
    T, E, B = make_alms(torch, lmax, batch, args.seed, dev)
    EB = [x for pair in zip(E, B) for x in pair]
    zeros = lambda: torch.zeros(npix, dtype=torch.float64, device=dev)  # noqa: E731
    map_t = [zeros() for _ in range(batch)]
    map_qu = [zeros() for _ in range(2 * batch)]
    out_t = [torch.zeros_like(T[0]) for _ in range(batch)]
    out_eb = [torch.zeros_like(T[0]) for _ in range(2 * batch)]
    stream = torch.cuda.current_stream()

    def step():
        S.execute_device(S.ALM2MAP, 0, T, map_t, geom, ainfo, batch, flags, stream)
        S.execute_device(S.ALM2MAP, 2, EB, map_qu, geom, ainfo, batch, flags, stream)
        S.execute_device(S.MAP2ALM, 0, out_t, map_t, geom, ainfo, batch, flags, stream)
        S.execute_device(S.MAP2ALM, 2, out_eb, map_qu, geom, ainfo, batch, flags, stream)

    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev).normal_()
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read + 1 GiB written")
        del src, dst
    log("warmup x%d" % args.warmup)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("timed x%d" % args.steps)
    S.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = S.profile_stats()
    S.profile(False)
    ms_per_step = 1e3 * dt / args.steps
    value = batch * args.steps / dt
    log("%.1f ms/step, %.3f TQU round trips/s" % (ms_per_step, value))
    for st in sorted(stats, key=lambda x: -x["ms"]):
        log("  %-24s %9.2f ms  %4d launches  %8.2f TFLOP/s  %8.1f GB/s" % (
            st["name"], st["ms"], st["launches"],
            st["flops"] / (st["ms"] * 1e-3) / 1e12 if st["ms"] > 0 else 0,
            st["bytes"] / (st["ms"] * 1e-3) / 1e9 if st["ms"] > 0 else 0))

    dom = max(stats, key=lambda x: x["ms"])
    avg_s = dom["ms"] * 1e-3 / dom["launches"]
    if dom["flops"] > 0:
        ach = dom["flops"] / dom["launches"] / avg_s
        roof = {"kernel": dom["name"], "bound": "mfma", "achieved": ach / 1e12,
                "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                "frac": ach / PEAK["fp64"], "traffic": None,
                "avg_launch_ms": avg_s * 1e3,
                "note": "FP64 work priced at the MI355X FP64 peak (vector and "
                        "matrix FP64 peaks are equal); flops = libsharp's flop "
                        "model per launch"}
    else:
        ach = dom["bytes"] / dom["launches"] / avg_s
        roof = {"kernel": dom["name"], "bound": "hbm", "achieved": ach / 1e9,
                "peak": PEAK["hbm"] / 1e9, "unit": "GB/s", "frac": ach / PEAK["hbm"],
                "traffic": None, "avg_launch_ms": avg_s * 1e3}

    cpu = None
    parity = None
    if not args.no_cpu:
        log("cpu baseline (reference libsharp, one realisation)")
        T0 = T[0].cpu().numpy()
        E0 = E[0].cpu().numpy()
        B0 = B[0].cpu().numpy()
        cpu, rmaps, ralms = cpu_baseline(nside, lmax, T0, E0, B0)
        if cpu is not None:
            gm = (map_t[0].cpu().numpy(), map_qu[0].cpu().numpy(), map_qu[1].cpu().numpy())
            ga = (out_t[0].cpu().numpy(), out_eb[0].cpu().numpy(), out_eb[1].cpu().numpy())
            rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))  # noqa: E731
            parity = {"maps_TQU_rel_l2": [rel(x, y) for x, y in zip(gm, rmaps)],
                      "alms_TEB_rel_l2": [rel(x, y) for x, y in zip(ga, ralms)]}
            log("parity vs reference:", parity)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TQU alm2map+map2alm round trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Gaussian a_lm, C_l^TT=1/(l(l+1)+1) x 5' beam, "
                "EE=0.1TT, BB=0.01TT, TE corr 0.5",
        "config": {"workload": "HEALPix Nside %d, lmax %d, TQU, batch %d, "
                               "alm2map+map2alm" % (nside, lmax, batch),
                   "nside": nside, "lmax": lmax, "batch": batch, "form": args.form},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if parity is not None:
        out["parity_vs_reference"] = parity
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

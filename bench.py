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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form mixed|fp64]
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
from sht_inputs import make_alms, spin0_alms  # noqa: E402

METRIC = ("alm2map+map2alm/s, Nside 2048 lmax 4096 TQU; mixed vs FP64 at 1e-6 "
          "rel err")
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "hbm": 8.0e12}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    """(model name, physical cores) of this host from /proc/cpuinfo."""
    model, cores = None, set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and core is not None:
                cores.add((phys, core))
                phys = core = None
        if core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return model, (len(cores) or os.cpu_count())


def cpu_baseline(nside, lmax, realisation, nreal, min_seconds=2.0, batched=True):
    """Reference libsharp (oracle/_ref, OpenMP) on the TQU workload, timed as
    its own benchmark does (libsharp/sharp_testsuite.c:554-577): one warm-up
    per job key (fills libsharp's nv timing oracle), then tries until at
    least two have run and min_seconds have accumulated; the best try is
    reported, with GFLOP/s from the opcnt libsharp returns.  One try = one
    realisation's alm2map (spin 0 + spin 2) + map2alm, ntrans = 1.  With
    `batched`, also one round trip of all nreal realisations as ntrans =
    nreal calls (libsharp shares the recurrence over them,
    libsharp/sharp_core.c:76-240).  Returns (dict, maps, alms) with the
    outputs of realisation 0, or (None, None, None) without the build.
    realisation(i) gives the (T, E, B) a_lm of realisation i (numpy)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.load_ref()
    if ref is None:
        return None, None, None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    model, phys = cpu_info()
    g = ref.healpix(nside)
    a = ref.triangular(lmax)
    npix = 12 * nside * nside
    # warm-up on a small job: the first call per (type, spin, ntrans) runs
    # libsharp's nv oracle
    gs = ref.healpix(16)
    as_ = ref.triangular(32)
    for nt in ((1, nreal) if batched else (1,)):
        for sp, na in ((0, 1), (2, 2)):
            ref.execute(O.ALM2MAP, sp, [np.zeros(561, np.complex128) for _ in range(na * nt)],
                        [np.zeros(3072) for _ in range(na * nt)], gs, as_, nt, O.DP)
            ref.execute(O.MAP2ALM, sp, [np.zeros(561, np.complex128) for _ in range(na * nt)],
                        [np.zeros(3072) for _ in range(na * nt)], gs, as_, nt, O.DP)

    def round_trip(alm_t, alm_eb, nt):
        map_t = [np.zeros(npix) for _ in range(nt)]
        map_qu = [np.zeros(npix) for _ in range(2 * nt)]
        out_t = [np.zeros_like(alm_t[0]) for _ in range(nt)]
        out_eb = [np.zeros_like(alm_t[0]) for _ in range(2 * nt)]
        t0 = time.perf_counter()
        ops = ref.execute(O.ALM2MAP, 0, alm_t, map_t, g, a, nt, O.DP)[1]
        ops += ref.execute(O.ALM2MAP, 2, alm_eb, map_qu, g, a, nt, O.DP)[1]
        ops += ref.execute(O.MAP2ALM, 0, out_t, map_t, g, a, nt, O.DP)[1]
        ops += ref.execute(O.MAP2ALM, 2, out_eb, map_qu, g, a, nt, O.DP)[1]
        return time.perf_counter() - t0, ops, (map_t, map_qu, out_t, out_eb)

    first, times, ops1 = None, [], 0
    while len(times) < min(2, nreal) or (sum(times) < min_seconds and len(times) < nreal):
        T0, E0, B0 = realisation(len(times))
        dt, ops1, outs = round_trip([T0], [E0, B0], 1)
        times.append(dt)
        if first is None:
            map_t, map_qu, out_t, out_eb = outs
            first = ((map_t[0], map_qu[0], map_qu[1]), (out_t[0], out_eb[0], out_eb[1]))
        log("  cpu: realisation %d, ntrans 1: %.2f s" % (len(times) - 1, dt))
    best = min(times)
    info = {"value": 1.0 / best, "unit": "TQU round trips/s", "cores": threads,
            "kind": "reference",
            "gflops": ops1 / best / 1e9,
            "cpu_model": model, "physical_cores_on_host": phys, "threads": threads,
            "sample": "libsharp alm2map spin 0+2 then map2alm spin 0+2 of one realisation "
                      "(ntrans 1) at Nside %d lmax %d, best of %d tries (%.1f s), "
                      "libsharp's own protocol" % (nside, lmax, len(times), sum(times))}
    if batched:
        rs = [realisation(i) for i in range(nreal)]
        dt, ops, _ = round_trip([r[0] for r in rs], [x for r in rs for x in (r[1], r[2])], nreal)
        log("  cpu: %d realisations as ntrans %d: %.2f s" % (nreal, nreal, dt))
        info["batched_ntrans%d" % nreal] = {
            "value": nreal / dt, "unit": "TQU round trips/s", "gflops": ops / dt / 1e9,
            "sample": "all %d realisations in one call per spin and direction (ntrans %d, "
                      "recurrence shared), one try (%.1f s)" % (nreal, nreal, dt)}
    return info, first[0], first[1]


def run_form(S, torch, step, warmup, steps):
    """W untimed steps, then K timed steps bracketed by device syncs, with the
    per-stage HIP events on; returns (seconds, stage stats)."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    S.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = S.profile_stats()
    S.profile(False)
    return dt, stats


def show(name, stats, ms_per_step, value):
    log("%s: %.1f ms/step, %.3f round trips/s" % (name, ms_per_step, value))
    for st in sorted(stats, key=lambda x: -x["ms"]):
        log("  %-24s %9.2f ms  %4d launches  %8.2f TFLOP/s  %8.1f GB/s" % (
            st["name"], st["ms"], st["launches"],
            st["flops"] / (st["ms"] * 1e-3) / 1e12 if st["ms"] > 0 else 0,
            st["bytes"] / (st["ms"] * 1e-3) / 1e9 if st["ms"] > 0 else 0))


def contraction_share(name, nt):
    """Share of the contraction in libsharp's per-triple flop model
    (libsharp/sharp_core_inc2.c:171-173,440-442,583-585)."""
    if "spin0" in name:
        return 4.0 * nt / (4 + 4.0 * nt)
    if "deriv1" in name:
        return 8.0 * nt / (12 + 8.0 * nt)
    return 16.0 * nt / (12 + 16.0 * nt)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(workload, stage):
    """HBM bytes per launch of a stage (FETCH_SIZE + WRITE_SIZE, calibrated
    against a device copy of known size), from the counter summary kept under
    profiles/; None when that workload/stage was not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    ent = tab.get("workloads", {}).get(workload, {}).get(stage)
    if ent is None:
        return None
    return ent["read_bytes"] + ent["write_bytes"]


def roofline(stats, mixed, nt, workload=None, bf16=False):
    """Roofline object of the slowest stage.  Legendre stages carry libsharp's
    flop model per launch (c0 + c1*ntrans per (m, ring pair, l) triple,
    libsharp/sharp_core_inc2.c:171-173,440-442); `achieved` is that
    algorithmic figure over the average launch time, priced at the peak of
    the unit the contraction runs on (BF16, FP32 or FP64 matrix cores).  Ring
    stages carry compulsory HBM bytes.  `traffic` is the stage's measured HBM
    bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, kept in
    profiles/hbm_traffic.json)."""
    dom = max(stats, key=lambda x: x["ms"])
    avg_s = dom["ms"] * 1e-3 / dom["launches"]
    traffic = measured_traffic(workload, dom["name"]) if workload else None
    common = {"kernel": dom["name"], "traffic": traffic, "avg_launch_ms": avg_s * 1e3,
              "traffic_source": "profiles/hbm_traffic.json" if traffic else None}
    if dom["flops"] > 0:
        ach = dom["flops"] / dom["launches"] / avg_s
        if mixed:
            c = contraction_share(dom["name"], nt)
            peak = PEAK["bf16"] if bf16 else PEAK["fp32"]
            out = dict(common, bound="mfma", achieved=ach / 1e12, peak=peak / 1e12,
                       unit="TFLOP/s", frac=ach / peak,
                       model_contraction_tflops=ach * c / 1e12,
                       fp64_recurrence_tflops=ach * (1 - c) / 1e12)
            if bf16:
                # the matrix cores run six BF16 piece products per FP32
                # product of the contraction: their issue rate, for reference
                out["bf16_mfma_issue_tflops"] = 6 * ach * c / 1e12
                out["bf16_mfma_issue_frac"] = 6 * ach * c / PEAK["bf16"]
                out["note"] = ("achieved = libsharp's flop model (whole launch) over the "
                               "dense BF16 MFMA peak; bf16_mfma_issue_* = the six BF16 "
                               "products per contraction product the matrix cores execute")
            else:
                out["note"] = ("achieved = libsharp's flop model (whole launch) over the "
                               "FP32 MFMA peak")
            return out
        return dict(common, bound="mfma", achieved=ach / 1e12, peak=PEAK["fp64"] / 1e12,
                    unit="TFLOP/s", frac=ach / PEAK["fp64"],
                    note="libsharp's flop model over the MI355X FP64 peak (vector and "
                         "matrix FP64 peaks are equal)")
    ach = dom["bytes"] / dom["launches"] / avg_s
    return dict(common, bound="hbm", achieved=ach / 1e9, peak=PEAK["hbm"] / 1e9,
                unit="GB/s", frac=ach / PEAK["hbm"])


def host_path(S, torch, T, EB, geom, ainfo, batch, npix, forms):
    """The drop-in sharp_execute on host (numpy) arrays: every call copies its
    inputs to HBM and its outputs back.  The best of three TQU round trips of
    the batch per form (the reference's best-of-N protocol,
    libsharp/sharp_testsuite.c:554-577), after one untimed round trip."""
    hT = [t.cpu().numpy() for t in T]
    hEB = [t.cpu().numpy() for t in EB]
    res = {}
    form_flags = {"mixed": S.MIXED_FP32, "bf16x3": S.MIXED_BF16X3, "fp64": 0}
    for form in forms:
        flags = S.DP | form_flags[form]
        map_t = [np.zeros(npix) for _ in range(batch)]
        map_qu = [np.zeros(npix) for _ in range(2 * batch)]
        out_t = [np.zeros_like(hT[0]) for _ in range(batch)]
        out_eb = [np.zeros_like(hT[0]) for _ in range(2 * batch)]

        def rt():
            S.execute(S.ALM2MAP, 0, hT, map_t, geom, ainfo, batch, flags)
            S.execute(S.ALM2MAP, 2, hEB, map_qu, geom, ainfo, batch, flags)
            S.execute(S.MAP2ALM, 0, out_t, map_t, geom, ainfo, batch, flags)
            S.execute(S.MAP2ALM, 2, out_eb, map_qu, geom, ainfo, batch, flags)
        rt()
        dt = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            rt()
            dt = min(dt, time.perf_counter() - t0)
        res[form] = {"value": batch / dt, "ms_per_step": dt * 1e3}
        log("host path %s: %.1f ms/step, %.3f round trips/s" % (form, dt * 1e3, batch / dt))
    res["unit"] = "TQU round trips/s"
    res["note"] = ("sharp_execute on host arrays: H2D of inputs and D2H of outputs "
                   "inside every call; best of 3 round trips")
    return res


def smoothing_workload(S, torch, dev, args):
    """smoothing_cxx's transform work on the TQU batch (Nside 2048, lmax 4096,
    8 realisations): map2alm_iter with 3 Jacobi iterations, then alm2map of
    the a_lm times a 5' Gaussian beam (the beam applied while the a_lm are
    gathered), for T (spin 0) and Q/U (spin 2), all on the GPU.  One step
    smooths the 8 TQU maps: 4 analyses + 4 syntheses per spin.  Both forms
    are timed; the error of the mixed form is stated against the FP64 form."""
    nside = args.nside
    lmax = args.lmax if args.lmax is not None else 2 * nside
    batch, niter = args.batch, 3
    fwhm = math.radians(5.0 / 60.0)
    npix = 12 * nside * nside
    geom = S.Geometry.healpix(nside)
    ainfo = S.AlmInfo.triangular(lmax)
    log("generating maps: nside %d lmax %d batch %d" % (nside, lmax, batch))
    T, E, B = make_alms(torch, lmax, batch, args.seed, dev)
    EB = [x for pair in zip(E, B) for x in pair]
    m_t = [torch.zeros(npix, dtype=torch.float64, device=dev) for _ in range(batch)]
    m_qu = [torch.zeros(npix, dtype=torch.float64, device=dev) for _ in range(2 * batch)]
    S.execute_device(S.ALM2MAP, 0, T, m_t, geom, ainfo, batch, S.DP)
    S.execute_device(S.ALM2MAP, 2, EB, m_qu, geom, ainfo, batch, S.DP)
    torch.cuda.synchronize()
    del T, E, B, EB
    res = {}
    for form in ("bf16x3", "fp64"):
        flags = S.DP | (S.MIXED_BF16X3 if form == "bf16x3" else 0)
        wt = [x.clone() for x in m_t]
        wqu = [x.clone() for x in m_qu]
        at = [torch.zeros(S.nalm_triangular(lmax), dtype=torch.complex128, device=dev)
              for _ in range(batch)]
        aeb = [torch.zeros_like(at[0]) for _ in range(2 * batch)]

        def step():
            # smoothing_cxx maps in place: restore the inputs first
            for d, x in zip(wt + wqu, m_t + m_qu):
                d.copy_(x)
            S.smoothing_device(wt, geom, ainfo, 0, fwhm, niter, at, flags)
            S.smoothing_device(wqu, geom, ainfo, 2, fwhm, niter, aeb, flags)

        log("%s: warmup x%d, timed x%d" % (form, args.warmup, args.steps))
        dt, stats = run_form(S, torch, step, args.warmup, args.steps)
        ms = 1e3 * dt / args.steps
        val = batch * args.steps / dt
        show(form, stats, ms, val)
        res[form] = {"value": val, "ms_per_step": ms, "maps": wt + wqu,
                     "roofline": roofline(stats, form != "fp64", batch,
                                          bf16=form == "bf16x3")}

    def rel(x, y):
        return float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y))
    err = max(rel(x, y) for x, y in zip(res["bf16x3"]["maps"], res["fp64"]["maps"]))
    out = {
        "metric": "smoothing_cxx TQU maps/s, Nside %d lmax %d, 3 iterations, 5' beam" % (nside, lmax),
        "value": res["bf16x3"]["value"], "unit": "TQU maps smoothed/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["bf16x3"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16-mfma contraction on 3-piece splits of f32, f64 recurrence/FFT",
        "data": "synthetic: TQU maps synthesised from the seeded a_lm of the headline workload",
        "config": {"workload": "smoothing_cxx core (map2alm_iter, smoothWithGauss, alm2map) "
                               "of 8 TQU maps", "nside": nside, "lmax": lmax,
                   "batch": batch, "iterations": niter, "fwhm_arcmin": 5.0},
        "roofline": res["bf16x3"]["roofline"],
        "fp64_form": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms_per_step"],
                      "roofline": res["fp64"]["roofline"]},
        "mixed_vs_fp64_max_rel_l2": err,
    }
    print(json.dumps(out), flush=True)


def batch_workload(S, torch, dev, args):
    """The uneven batch: 24 spin-0 maps each at Nside 128, 256, 512, 1024
    (lmax = 2 Nside) and 8 Gauss-Legendre grids each at lmax 255 and 767
    (nlat = lmax+1, nlon = 2 lmax+1); a_lm with C_l = 1/(l(l+1)+1).  One step
    is alm2map + map2alm of all 112 maps.  Each form runs as two batches (one
    persistent Legendre launch per direction, tile shape tuned at creation,
    replayed HIP graphs): BF16-split matrix cores (`value`), FP32 matrix
    cores and FP64; the FP64 form one call per geometry is timed beside them.
    The error of every map and round-trip a_lm against the FP64 per-geometry
    form is reported."""
    specs = [("healpix", n, 2 * n, 24) for n in (128, 256, 512, 1024)]
    specs += [("gauss", 255, 255, 8), ("gauss", 767, 767, 8)]
    gen = torch.Generator(device=dev).manual_seed(args.seed)
    jobs = []
    for kind, n, lmax, nt in specs:
        geom = S.Geometry.healpix(n) if kind == "healpix" else S.Geometry.iso("gauss", lmax + 1, 2 * lmax + 1)
        ainfo = S.AlmInfo.triangular(lmax)
        alms = spin0_alms(torch, lmax, nt, gen, dev)
        npix = geom.npix
        jobs.append({"name": "%s %d lmax %d x%d" % (kind, n, lmax, nt), "geom": geom,
                     "ainfo": ainfo, "nt": nt, "alms": alms,
                     "maps64": [torch.zeros(npix, dtype=torch.float64, device=dev) for _ in range(nt)],
                     "out64": [torch.zeros_like(alms[0]) for _ in range(nt)]})
    nmaps = sum(j["nt"] for j in jobs)

    def step_fp64():
        for j in jobs:
            S.execute_device(S.ALM2MAP, 0, j["alms"], j["maps64"], j["geom"], j["ainfo"], j["nt"], S.DP)
            S.execute_device(S.MAP2ALM, 0, j["out64"], j["maps64"], j["geom"], j["ainfo"], j["nt"], S.DP)

    dt64, stats64 = run_form(S, torch, step_fp64, args.warmup, args.steps)
    ms64 = 1e3 * dt64 / args.steps
    show("per-geometry fp64", stats64, ms64, nmaps / (dt64 / args.steps))

    def rel(x, y):
        return float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y))
    res = {}
    for form, flag in (("bf16x3", S.MIXED_BF16X3), ("fp32", S.MIXED_FP32), ("fp64", 0)):
        for j in jobs:
            j["maps"] = [torch.zeros_like(x) for x in j["maps64"]]
            j["out"] = [torch.zeros_like(x) for x in j["out64"]]
        log("batch %s: creating plans (autotuning, graph capture)" % form)
        bs = S.Batch([(S.ALM2MAP, 0, j["alms"], j["maps"], j["geom"], j["ainfo"], j["nt"]) for j in jobs],
                     S.DP | flag, use_graph=True)
        ba = S.Batch([(S.MAP2ALM, 0, j["out"], j["maps"], j["geom"], j["ainfo"], j["nt"]) for j in jobs],
                     S.DP | flag, use_graph=True)

        def step():
            bs.execute()
            ba.execute()

        dt, stats = run_form(S, torch, step, args.warmup, args.steps)
        ms = 1e3 * dt / args.steps
        show("batch " + form, stats, ms, nmaps / (dt / args.steps))
        per = {}
        for j in jobs:
            per[j["name"]] = {
                "maps_max_rel_l2": max(rel(a, b) for a, b in zip(j["maps"], j["maps64"])),
                "roundtrip_alm_max_rel_l2": max(rel(a, b) for a, b in zip(j["out"], j["out64"]))}
        res[form] = {"value": nmaps / (dt / args.steps), "ms_per_step": ms,
                     "tiles": {"synthesis": bs.config(), "analysis": ba.config()},
                     "error_vs_fp64_per_geometry": per}
        log("batch %s: per-map error vs FP64 per geometry:" % form, per)
        bs.close()
        ba.close()
    out = {
        "metric": "uneven batch alm2map+map2alm, maps/s (BF16-split batched form)",
        "value": res["bf16x3"]["value"], "unit": "map round trips/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["bf16x3"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16-mfma contraction on 3-piece splits of f32 (spin-0 synthesis on f32 "
                 "mfma), f64 recurrence/FFT",
        "data": "synthetic: seeded Gaussian a_lm, C_l=1/(l(l+1)+1)",
        "config": {"workload": "96 HEALPix maps (24 each Nside 128/256/512/1024, lmax=2 Nside) "
                               "+ 16 Gauss-Legendre grids (8 each lmax 255/767), spin 0"},
        "forms": res,
        "fp64_per_geometry": {"value": nmaps / (dt64 / args.steps), "ms_per_step": ms64},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--form", default="bf16x3", choices=["mixed", "bf16x3", "fp64"],
                    help="form whose throughput is `value` (mixed: FP32 matrix "
                         "cores; bf16x3: BF16 matrix cores on three-piece splits); "
                         "the other ones are timed beside it")
    ap.add_argument("--only", action="store_true",
                    help="time only --form (no comparison run)")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-batched", action="store_true",
                    help="skip the reference's ntrans=batch CPU figure")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (calibrates HBM counters)")
    ap.add_argument("--streams", type=int, default=1, choices=[1, 2],
                    help="streams for the T and Q/U transforms of a step")
    ap.add_argument("--no-host-path", dest="host_path", action="store_false",
                    help="skip timing sharp_execute on host arrays (H2D/D2H copies "
                         "included; reported beside value as host_path, never as value)")
    ap.add_argument("--workload", default="tqu", choices=["tqu", "batch", "smoothing"],
                    help="tqu: the headline workload; batch: the uneven batch "
                         "(96 HEALPix maps of four sizes + 16 Gauss-Legendre grids); "
                         "smoothing: smoothing_cxx of the TQU batch (3 Jacobi "
                         "iterations, 5' beam)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the multi-GPU form of this path "
              "(realisations split over GPUs, or an m-sharded transform with one "
              "all-to-all of the phases per direction) is not built; use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import sharp_hip as S

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload == "batch":
        batch_workload(S, torch, dev, args)
        return
    if args.workload == "smoothing":
        smoothing_workload(S, torch, dev, args)
        return
    nside = args.nside
    lmax = args.lmax if args.lmax is not None else 2 * nside
    batch = args.batch
    npix = 12 * nside * nside
    geom = S.Geometry.healpix(nside)
    ainfo = S.AlmInfo.triangular(lmax)

    log("generating inputs: nside %d lmax %d batch %d" % (nside, lmax, batch))
    T, E, B = make_alms(torch, lmax, batch, args.seed, dev)
    EB = [x for pair in zip(E, B) for x in pair]
    zeros = lambda: torch.zeros(npix, dtype=torch.float64, device=dev)  # noqa: E731
    stream = torch.cuda.current_stream()
    # default-priority side streams, ordered after the input generation
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for x in streams:
        x.wait_stream(stream)

    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev).normal_()
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read + 1 GiB written")
        del src, dst

    forms = [args.form] if args.only else \
        [args.form] + [f for f in ("mixed", "bf16x3", "fp64") if f != args.form]
    form_flags = {"mixed": S.MIXED_FP32, "bf16x3": S.MIXED_BF16X3, "fp64": 0}
    res = {}
    for form in forms:
        flags = S.DP | form_flags[form]
        map_t = [zeros() for _ in range(batch)]
        map_qu = [zeros() for _ in range(2 * batch)]
        out_t = [torch.zeros_like(T[0]) for _ in range(batch)]
        out_eb = [torch.zeros_like(T[0]) for _ in range(2 * batch)]

        # T (spin 0) and Q/U (spin 2) are independent: with --streams 2 they
        # are enqueued on two streams (the library keeps work buffers per
        # stream), so the kernels of one overlap those of the other
        st_t, st_qu = (stream, stream) if args.streams == 1 else (streams[0], streams[1])

        def step():
            S.execute_device(S.ALM2MAP, 0, T, map_t, geom, ainfo, batch, flags, st_t)
            S.execute_device(S.ALM2MAP, 2, EB, map_qu, geom, ainfo, batch, flags, st_qu)
            S.execute_device(S.MAP2ALM, 0, out_t, map_t, geom, ainfo, batch, flags, st_t)
            S.execute_device(S.MAP2ALM, 2, out_eb, map_qu, geom, ainfo, batch, flags, st_qu)

        log("%s: warmup x%d, timed x%d" % (form, args.warmup, args.steps))
        dt, stats = run_form(S, torch, step, args.warmup, args.steps)
        ms = 1e3 * dt / args.steps
        val = batch * args.steps / dt
        show(form, stats, ms, val)
        res[form] = {"value": val, "ms_per_step": ms, "stats": stats,
                     "roofline": roofline(stats, form != "fp64", batch,
                                          "nside%d_lmax%d_tqu%d_%s" % (nside, lmax, batch, form),
                                          bf16=form == "bf16x3"),
                     "maps": map_t + map_qu, "alms": out_t + out_eb}

    # mixed forms vs FP64 on the GPU: maps after alm2map, a_lm after the
    # round trip
    def rel(x, y):
        return float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y))

    errs = {}
    for form in res:
        if form == "fp64" or "fp64" not in res:
            continue
        mm, fm = res[form]["maps"], res["fp64"]["maps"]
        ma, fa = res[form]["alms"], res["fp64"]["alms"]
        e = {
            "maps_T_max_rel_l2": max(rel(mm[i], fm[i]) for i in range(batch)),
            "maps_QU_max_rel_l2": max(rel(mm[batch + i], fm[batch + i])
                                      for i in range(2 * batch)),
            "roundtrip_alm_T_max_rel_l2": max(rel(ma[i], fa[i]) for i in range(batch)),
            "roundtrip_alm_EB_max_rel_l2": max(rel(ma[batch + i], fa[batch + i])
                                               for i in range(2 * batch)),
        }
        e["max"] = max(e.values())
        e["bound"] = 1e-6
        errs[form] = e
        log("%s vs FP64:" % form, e)
    mix_err = errs.get(args.form)

    host = None
    if args.host_path:
        host = host_path(S, torch, T, EB, geom, ainfo, batch, npix, forms)

    cpu = None
    parity = None
    if not args.no_cpu and "fp64" in res:
        log("cpu baseline (reference libsharp, one realisation)")
        cpu, rmaps, ralms = cpu_baseline(
            nside, lmax, lambda i: (T[i].cpu().numpy(), E[i].cpu().numpy(),
                                    B[i].cpu().numpy()), batch,
            batched=not args.no_cpu_batched)
        if cpu is not None:
            fm, fa = res["fp64"]["maps"], res["fp64"]["alms"]
            gm = (fm[0].cpu().numpy(), fm[batch].cpu().numpy(), fm[batch + 1].cpu().numpy())
            ga = (fa[0].cpu().numpy(), fa[batch].cpu().numpy(), fa[batch + 1].cpu().numpy())
            rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))  # noqa: E731
            parity = {"maps_TQU_rel_l2": [rel(x, y) for x, y in zip(gm, rmaps)],
                      "alms_TEB_rel_l2": [rel(x, y) for x, y in zip(ga, ralms)]}
            log("FP64 form vs reference:", parity)

    main_form = res[args.form]
    out = {
        "metric": METRIC,
        "value": main_form["value"],
        "unit": "TQU alm2map+map2alm round trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_form["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f32-mfma contraction, f64 recurrence/FFT",
                  "bf16x3": "bf16-mfma contraction on 3-piece splits of f32 (6 products, "
                            "f32 accumulation; the spin-0 synthesis and the exponent-"
                            "scaled spin-2 synthesis blocks on f32 mfma), f64 "
                            "recurrence/FFT",
                  "fp64": "f64"}[args.form],
        "data": "synthetic: seeded Gaussian a_lm, C_l^TT=1/(l(l+1)+1) x 5' beam, "
                "EE=0.1TT, BB=0.01TT, TE corr 0.5",
        "config": {"workload": "HEALPix Nside %d, lmax %d, TQU, batch %d, "
                               "alm2map+map2alm" % (nside, lmax, batch),
                   "nside": nside, "lmax": lmax, "batch": batch, "form": args.form,
                   "streams": args.streams},
        "roofline": main_form["roofline"],
        "cpu_baseline": cpu,
    }
    for form in res:
        if form != args.form:
            out[form + "_form"] = {"value": res[form]["value"],
                                   "ms_per_step": res[form]["ms_per_step"],
                                   "roofline": res[form]["roofline"]}
            if form in errs:
                out[form + "_form"]["vs_fp64"] = errs[form]
    if mix_err is not None:
        out["mixed_vs_fp64"] = mix_err
    if host is not None:
        out["host_path"] = host
    if parity is not None:
        out["fp64_vs_reference"] = parity
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

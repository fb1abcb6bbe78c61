#!/usr/bin/env python3
"""bench.py -- pixel-geodesics/s of the per-pixel ray tracer on one MI355X.

One step = one render of the 1024x1024 image of a 256x256 HARM torus dump at
230 GHz (inclination 60 deg, 40 M field of view; FP64 geodesics, FP32
transfer): every pixel's null geodesic plus its radiative transfer.  Inputs
are resident in HBM before the timed region; outputs stay in HBM.

Prints one JSON line (rank 0).  Extra fields:
  roofline      algorithmic FLOP rate of the render's kernels (approach and
                trace) vs the FP64/FP32 vector peak for its mix of work
  grad          image + d/d(ln M_unit, R_high) throughput beside plain images
  cpu_baseline  the FP64 CPU oracle (OpenMP, all host cores) on a pixel sample
  parity        GPU vs oracle on that same sample

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--cpu-sample N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
import warnings

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixel-geodesics/s at 1024^2, 230 GHz; image+grad/s vs plain, <=1e-4 flux err"
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (dense) peak
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector (dense) peak

WORKLOAD = dict(workload="harm2d-torus-256x256_img1024x1024_230GHz_i60", dump="256x256x1",
                image="1024x1024", freq_hz=2.3e11, inclination_deg=60.0, M_unit=1e23,
                R_high=10.0, R_low=1.0, stepsize=0.03, fov_M=40.0)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def _parity(g, ora, gate):
    """GPU pixels g against the oracle's ora (same shape); gate: the oracle's
    distance of tau from the ln 1000 gate at its closest sample."""
    floor = 1e-6 * np.max(np.abs(ora))
    r = np.abs(g - ora) / np.maximum(np.abs(ora), floor)
    near = gate < 1e-4
    return {
        "pixels": int(g.shape[0]), "flux_rel": float(g.sum() / ora.sum() - 1.0),
        "max_pix_rel": float(r.max()), "n_pix_over_1e-3": int((r > 1e-3).sum()),
        "floor": "1e-6*max(I)",
        "gate_adjacent": int(near.sum()),
        "max_pix_rel_off_gate": float(r[~near].max()) if (~near).any() else 0.0,
        "gate_note": "gate_adjacent: pixels whose FP64 tau passes within 1e-4 of the "
                     "ln 1000 gate at a sample, where the FP32 tau may add or drop one "
                     "exp(-6.9)-weighted sample",
    }


def _sample(npix, n):
    stride = max(1, npix // n)
    return np.arange(0, npix, stride, dtype=np.int32)[:n], stride


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grad-steps", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20,
                    help="pixels of the workload traced by the CPU oracle (default: all "
                         "of the 1024^2 image, about 10 s on 16 cores; 0 = skip)")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--emission", choices=["thsynch", "thsynchav", "kappa"], default="thsynch",
                    help="emission model of the c1 image (default: thsynch, the path's); "
                         "the others have no gradient leg and no roofline (the frozen "
                         "operation counts are the thermal fit's)")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4"], default="c1",
                    help="c1: the headline 1024^2 image (default); c2: 6 inclinations x 8 "
                         "frequencies at 512^2, one launch vs 48; c3: 3D 192x96x96 dump at "
                         "1024^2; c4: image+gradient on the 3D dump at 512^2")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(pixels shard trivially; see the sharding design). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from raptor_amd import flops, rt, torus

    if args.workload == "c2":
        return bench_c2(args)
    if args.workload in ("c3", "c4"):
        return bench_3d(args)

    W = args.size
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dump = torus.fishbone_moncrief(256, 256)
    emission = {"thsynch": rt.EMIS_THSYNCH, "thsynchav": rt.EMIS_THSYNCHAV,
                "kappa": rt.EMIS_KAPPA}[args.emission]
    thermal = emission == rt.EMIS_THSYNCH
    tr = rt.Tracer(0)
    tr.set_params(width=W, height=W, M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                  R_low=WORKLOAD["R_low"], stepsize=WORKLOAD["stepsize"],
                  cam_size=WORKLOAD["fov_M"], emission=emission)
    tr.load_dump(dump)
    incl = [WORKLOAD["inclination_deg"]]
    freq = [WORKLOAD["freq_hz"]]
    npix = W * W
    d_img = tr.alloc(npix * 8)
    d_dimg = tr.alloc(npix * 16)
    d_cnt = tr.alloc(npix * 8)

    # counters pass (untimed): per-ray steps and transfer samples
    tr.render_device(incl, freq, d_img, None, d_cnt)
    tr.synchronize()
    cnt = tr.d2h(d_cnt, (npix, 2), np.int32)
    f64, f32 = flops.image_flops_split(cnt[:, 0], cnt[:, 1], 1)
    alg_flops = f64 + f32
    alg_flops_grad = flops.image_flops(cnt[:, 0], cnt[:, 1], 1, grad=True)
    peak = flops.mix_peak(f64, f32, FP64_PEAK_TFLOPS, FP32_PEAK_TFLOPS)

    for _ in range(args.warmup):
        tr.render_device(incl, freq, d_img)
    tr.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tr.render_device(incl, freq, d_img)
        tr.synchronize()
        kms.append(tr.last_kernel_ms())
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = npix * args.steps / wall
    kern_ms = float(np.mean(kms))
    achieved = alg_flops / (kern_ms * 1e-3) / 1e12

    # gradient leg: image + its two parameter derivatives (thermal fit only)
    gms = [float("nan")]
    grad_rate = float("nan")
    if thermal and args.grad_steps > 0:
        tr.render_device(incl, freq, d_img, d_dimg)
        tr.synchronize()
        tg0 = time.perf_counter()
        gms = []
        for _ in range(args.grad_steps):
            tr.render_device(incl, freq, d_img, d_dimg)
            tr.synchronize()
            gms.append(tr.last_kernel_ms())
        tg1 = time.perf_counter()
        grad_rate = args.grad_steps / (tg1 - tg0)
        # the image the parity check reads is the plain render's
        tr.render_device(incl, freq, d_img)
        tr.synchronize()
    img = tr.d2h(d_img, (npix,), np.float64)
    jy = tr.jansky_factor()

    # the library's default launch: approach kernel for all rays of launches
    # of at least 2^20 rays, then the persistent kernel resuming them
    launch = tr.get_launch_config()
    split = launch["approach"] > 0 or (launch["approach"] == 0 and npix >= 1 << 20)
    kernels = ("approach_kernel<0> + trace_kernel_3w<1,false,false,0,true,0>" if split
               else "trace_kernel_3w<1,false,false,0,false,0>")
    traffic = issue = None
    tfile = os.path.join(ROOT, "profiles", "trace_kernel_traffic.json")
    if W == 1024 and split and os.path.exists(tfile):
        with open(tfile) as fh:
            prof = json.load(fh)
        traffic = prof.get("bytes_per_launch")
        insts = prof.get("valu_wave_insts_per_launch")
        if insts:
            # a wave64 VALU instruction occupies its SIMD for 4 cycles
            issue = insts * 4 / (kern_ms * 1e-3 * 1024 * 2.4e9)

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pixel-geodesics/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (seeded Fishbone-Moncrief torus, HARM MKS, a=0.9375)",
        "config": dict(WORKLOAD, image="%dx%d" % (W, W), emission=args.emission),
        "launch": dict(launch, approach_effective=100 if split else -1),
        "images_per_s": 1e3 / ms_per_step,
        "flux_jy": float(img.sum() * jy),
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_note": "HBM bytes per render (both kernels) from rocprofv3 "
                            "FETCH_SIZE/WRITE_SIZE, copy-calibrated "
                            "(profiles/trace_kernel_traffic.json)",
            "kernel": kernels,
            "kernel_note": "the render's kernels, timed together with HIP events on the "
                           "library stream; the roofline covers their sum",
            "kernel_ms": kern_ms,
            "valu_issue_frac": issue,
            "valu_issue_note": "wave64 VALU instructions per render (rocprofv3 SQ_INSTS_VALU, "
                               "profiles/trace_kernel_traffic.json) x 4 cycles over 1024 SIMDs "
                               "x 2.4 GHz x kernel_ms: the share of issue slots used",
            "alg_flops_per_launch": alg_flops,
            "alg_flops_fp64": f64,
            "alg_flops_fp32": f32,
            "peak_note": "FP64/FP32 mix-weighted vector peak (78.6 / 157.3 TFLOP/s)",
            "rk2_steps_per_launch": int(cnt[:, 0].sum()),
            "samples_per_launch": int(cnt[:, 1].sum()),
        },
        "grad": {
            "image_plus_grad_per_s": grad_rate,
            "image_per_s": 1e3 / ms_per_step,
            "ratio": grad_rate / (1e3 / ms_per_step),
            "kernel_ms": float(np.mean(gms)),
            "achieved_tflops": alg_flops_grad / (np.mean(gms) * 1e-3) / 1e12,
        },
    }
    if not thermal:
        result["roofline"] = None
        result["roofline_note"] = ("no roofline: the frozen operation counts "
                                   "(raptor_amd/flops.py) are the thermal fit's")
        result["grad"] = None

    if args.cpu_sample > 0:
        from oracle import oracle as O
        n = min(args.cpu_sample, npix)
        stride = max(1, npix // n)
        pix = np.arange(0, npix, stride, dtype=np.int32)[:n]
        P = O.params(dump.grid, width=W, height=W, inclination=incl[0],
                     M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                     R_low=WORKLOAD["R_low"], cam_size=WORKLOAD["fov_M"],
                     stepsize=WORKLOAD["stepsize"], emission=emission)
        cores = host_cores()
        c0 = time.perf_counter()
        ora, gate = O.render(P, dump.prims, freq, pixels=pix, gate=True, nthreads=cores)
        ora, gate = ora[:, 0], gate[:, 0]
        c1 = time.perf_counter()
        cpu_rate = n / (c1 - c0)
        result["cpu_baseline"] = {
            "value": cpu_rate, "unit": "pixel-geodesics/s", "cores": cores, "kind": "port",
            "sample": "%d of %d pixels (every %d-th) of the same workload, FP64 oracle, "
                      "OpenMP dynamic schedule, %.1f s wall" % (n, npix, stride, c1 - c0)
                      if n < npix else
                      "all %d pixels of the same workload, FP64 oracle, OpenMP dynamic "
                      "schedule, %.1f s wall" % (npix, c1 - c0),
        }
        result["speedup_vs_cpu"] = value / cpu_rate
        result["parity"] = _parity(img[pix], ora, gate)

    tr.free(d_img)
    tr.free(d_dimg)
    tr.free(d_cnt)
    tr.close()
    print(json.dumps(result), flush=True)


def _roofline(tr, incl, freqs, kernel_ms, grad=False):
    """Algorithmic FLOP rate of a render whose kernels took kernel_ms, from a
    counters pass (per-ray steps and samples) of the same launch."""
    from raptor_amd import flops
    n = len(incl) * tr.width * tr.height
    d_img = tr.alloc(n * len(freqs) * 8)
    d_dimg = tr.alloc(n * len(freqs) * 16) if grad else None
    d_cnt = tr.alloc(n * 8)
    try:
        tr.render_device(incl, freqs, d_img, d_dimg, d_cnt)
        tr.synchronize()
        cnt = tr.d2h(d_cnt, (n, 2), np.int32)
    finally:
        tr.free(d_img)
        if d_dimg:
            tr.free(d_dimg)
        tr.free(d_cnt)
    f64, f32 = flops.image_flops_split(cnt[:, 0], cnt[:, 1], len(freqs), grad=grad)
    peak = flops.mix_peak(f64, f32, FP64_PEAK_TFLOPS, FP32_PEAK_TFLOPS)
    achieved = (f64 + f32) / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "alg_flops_per_launch": f64 + f32,
            "rk2_steps_per_launch": int(cnt[:, 0].sum()),
            "samples_per_launch": int(cnt[:, 1].sum()), "kernel_ms": kernel_ms}


def _timed(tr, fn, steps):
    """Mean kernel ms (HIP events) and wall ms per call of fn()."""
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
        tr.synchronize()
        kms.append(tr.last_kernel_ms())
    return float(np.mean(kms)), 1e3 * (time.perf_counter() - t0) / steps


def bench_c2(args):
    """Batched: 6 inclinations x 8 log-spaced frequencies (86-690 GHz) at 512^2
    in one persistent launch (autotuned shape) against the same 48 images as
    single-image launches replayed from one HIP graph."""
    from raptor_amd import autotune, rt, torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dump = torus.fishbone_moncrief(256, 256)
    W = 512
    incl = [10.0, 25.0, 40.0, 55.0, 70.0, 85.0]
    freqs = list(86e9 * (690.0 / 86.0) ** (np.arange(8) / 7.0))
    tr = rt.Tracer(0)
    tr.set_params(width=W, height=W, M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                  R_low=WORKLOAD["R_low"])
    tr.load_dump(dump)
    npix = W * W
    tune = autotune.tune(tr, incl, freqs, reps=1, path=None)
    d_b = tr.alloc(len(incl) * npix * len(freqs) * 8)
    d_s = tr.alloc(len(incl) * len(freqs) * npix * 8)
    for _ in range(max(1, args.warmup)):
        tr.render_device(incl, freqs, d_b)
    tr.synchronize()
    kb, wb = _timed(tr, lambda: tr.render_device(incl, freqs, d_b), args.steps)
    # 48 single-image launches, captured once and replayed
    tr.set_launch_config()
    tr.capture_begin()
    k = 0
    for inc in incl:
        for nu in freqs:
            tr.render_device([inc], [nu], d_s + k * npix * 8)
            k += 1
    g = tr.capture_end()
    g.launch()
    tr.synchronize()
    ks, ws = _timed(tr, g.launch, args.steps)
    b = tr.d2h(d_b, (len(incl), npix, len(freqs)), np.float64)
    s = tr.d2h(d_s, (len(incl), len(freqs), npix), np.float64)
    same = bool(np.array_equal(b, np.transpose(s, (0, 2, 1))))
    tr.set_launch_config(**tune["config"])
    roof = _roofline(tr, incl, freqs, kb)
    g.destroy()
    tr.free(d_b)
    tr.free(d_s)
    tr.close()
    rays = len(incl) * npix
    cpu = parity = None
    if args.cpu_sample > 0:
        # the FP64 oracle on every 64th pixel of each inclination, all 8
        # frequencies: the CPU baseline, and parity for all 48 images
        from oracle import oracle as O
        pix, stride = _sample(npix, min(npix // 64, args.cpu_sample))
        cores = host_cores()
        oras, gates = [], []
        c0 = time.perf_counter()
        for inc in incl:
            P = O.params(dump.grid, width=W, height=W, inclination=inc,
                         M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                         R_low=WORKLOAD["R_low"])
            o, gt = O.render(P, dump.prims, freqs, pixels=pix, gate=True, nthreads=cores)
            oras.append(o)
            gates.append(gt)
        c1 = time.perf_counter()
        n = len(incl) * pix.size
        cpu = {"value": n / (c1 - c0), "unit": "pixel-geodesics/s", "cores": cores,
               "kind": "port",
               "sample": "every %d-th pixel of each of the %d inclinations (%d rays, %d "
                         "frequencies each), FP64 oracle, OpenMP dynamic schedule, %.1f s "
                         "wall" % (stride, len(incl), n, len(freqs), c1 - c0)}
        parity = _parity(b[:, pix, :].reshape(-1), np.stack(oras).reshape(-1),
                         np.stack(gates).reshape(-1))
        parity["images"] = len(incl) * len(freqs)
        parity["max_image_flux_rel"] = float(np.max(np.abs(
            b[:, pix, :].sum(axis=1) / np.stack(oras).sum(axis=1) - 1.0)))
    print(json.dumps({
        "metric": METRIC, "workload": "c2", "unit": "pixel-geodesics/s",
        "value": rays / (wb * 1e-3), "ms_per_step": wb, "steps": args.steps,
        "config": dict(workload="harm2d-torus-256x256_img512x512_6incl_8freq", dump="256x256x1",
                       image="512x512", inclinations=incl, freqs_hz=freqs),
        "batched": {"kernel_ms": kb, "wall_ms": wb, "launch_config": tune["config"],
                    "images_per_s": len(incl) * len(freqs) / (wb * 1e-3), "roofline": roof},
        "per_image_graph": {"launches": len(incl) * len(freqs), "kernel_ms": ks,
                            "wall_ms": ws,
                            "images_per_s": len(incl) * len(freqs) / (ws * 1e-3)},
        "speedup_batched": ws / wb, "batched_equals_per_image": same,
        "cpu_baseline": cpu, "parity": parity,
        "speedup_vs_cpu": rays / (wb * 1e-3) / cpu["value"] if cpu else None,
        "autotune_table": [[c, round(m, 3)] for c, m in tune["table"]],
    }), flush=True)


def bench_3d(args):
    """c3: the 3D 192x96x96 torus (m=1..4 azimuthal modes, 10%) at 1024^2, 230
    GHz.  c4: the same dump at 512^2, image + d/d(ln M_unit, R_high) against
    the image alone, with a central-difference check of the gradient."""
    from raptor_amd import rt, torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dump = torus.fishbone_moncrief(192, 96, 96)
    W = 1024 if args.workload == "c3" else 512
    npix = W * W
    incl, freq = [WORKLOAD["inclination_deg"]], [WORKLOAD["freq_hz"]]
    tr = rt.Tracer(0)
    kw = dict(width=W, height=W, R_low=WORKLOAD["R_low"])
    tr.set_params(M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"], **kw)
    tr.load_dump(dump)
    d_img = tr.alloc(npix * 8)
    d_dimg = tr.alloc(npix * 16)
    for _ in range(max(1, args.warmup)):
        tr.render_device(incl, freq, d_img)
    tr.synchronize()
    ki, wi = _timed(tr, lambda: tr.render_device(incl, freq, d_img), args.steps)
    out = {"metric": METRIC, "workload": args.workload, "unit": "pixel-geodesics/s",
           "value": npix / (wi * 1e-3), "ms_per_step": wi, "steps": args.steps,
           "config": dict(workload="harm3d-torus-192x96x96_img%dx%d_230GHz_i60" % (W, W),
                          dump="192x96x96", image="%dx%d" % (W, W),
                          prims_MB=dump.prims.size * 4 / 1e6),
           "image": {"kernel_ms": ki, "wall_ms": wi, "images_per_s": 1e3 / wi,
                     "roofline": _roofline(tr, incl, freq, ki)}}
    img = tr.d2h(d_img, (npix,), np.float64)
    if args.cpu_sample > 0:
        # the FP64 oracle on a pixel sample: the CPU baseline and parity
        from oracle import oracle as O
        pix, stride = _sample(npix, min(npix // 64, args.cpu_sample))
        P = O.params(dump.grid, width=W, height=W, inclination=incl[0],
                     M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                     R_low=WORKLOAD["R_low"])
        cores = host_cores()
        c0 = time.perf_counter()
        ora0, gate0 = O.render(P, dump.prims, freq, pixels=pix, gate=True, nthreads=cores)
        c1 = time.perf_counter()
        out["cpu_baseline"] = {
            "value": pix.size / (c1 - c0), "unit": "pixel-geodesics/s", "cores": cores,
            "kind": "port",
            "sample": "every %d-th pixel (%d rays), FP64 oracle, OpenMP dynamic schedule, "
                      "%.1f s wall" % (stride, pix.size, c1 - c0)}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["parity"] = _parity(img[pix], ora0[:, 0], gate0[:, 0])
    if args.workload == "c4":
        tr.render_device(incl, freq, d_img, d_dimg)
        tr.synchronize()
        kg, wg = _timed(tr, lambda: tr.render_device(incl, freq, d_img, d_dimg), args.steps)
        an = tr.d2h(d_dimg, (npix, 2), np.float64)
        out["grad"] = {"kernel_ms": kg, "wall_ms": wg, "image_plus_grad_per_s": 1e3 / wg,
                       "ratio_vs_image": wi / wg,
                       "roofline": _roofline(tr, incl, freq, kg, grad=True)}
    if args.workload == "c4" and args.cpu_sample > 0:
        # the CPU leg's check of the gradient: central differences (step 1e-3
        # and 5e-4) of the FP64 oracle on a pixel sample; GPU-image
        # differences would be FP32-noise limited
        M0, R0 = WORKLOAD["M_unit"], WORKLOAD["R_high"]
        pix = np.arange(0, npix, max(1, npix // 2048), dtype=np.int32)[:2048]

        def ora(M, R):
            P = O.params(dump.grid, width=W, height=W, inclination=incl[0], M_unit=M,
                         R_high=R, R_low=WORKLOAD["R_low"])
            return O.render(P, dump.prims, freq, pixels=pix, nthreads=host_cores())[:, 0]

        stats = {}
        for q, name in enumerate(("dlnMunit", "dRhigh")):
            fds = []
            for h in (1e-3, 5e-4):
                if q == 0:
                    fd = (ora(M0 * np.exp(h), R0) - ora(M0 * np.exp(-h), R0)) / (2 * h)
                else:
                    fd = (ora(M0, R0 * (1 + h)) - ora(M0, R0 * (1 - h))) / (2 * h * R0)
                fds.append(fd)
            a = an[pix, q]
            den = np.maximum(np.abs(a), 1e-3 * np.abs(a).max())
            smooth = np.abs(fds[0] - fds[1]) / den < 1e-3
            err = np.abs(fds[0] - a) / den
            stats[name] = {"pixels": int(pix.size), "smooth_frac": float(smooth.mean()),
                           "max_rel_smooth": float(err[smooth].max()),
                           "frac_within_1e-3": float((err[smooth] < 1e-3).mean()),
                           "sum_rel": float(fds[0][smooth].sum() / a[smooth].sum() - 1)}
        out["grad"]["fd_check"] = stats
        out["grad"]["fd_note"] = ("GPU tangents vs central differences (step 1e-3) of the "
                                  "FP64 oracle on 2048 sampled pixels; pixels whose quotient "
                                  "changes between steps 1e-3 and 5e-4 (gate flips) excluded")
    tr.free(d_img)
    tr.free(d_dimg)
    tr.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

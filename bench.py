#!/usr/bin/env python3
"""bench.py -- pixel-geodesics/s of the per-pixel ray tracer on one MI355X.

One step = one render of the 1024x1024 image of a 256x256 HARM torus dump at
230 GHz (inclination 60 deg, 40 M field of view; FP64 geodesics, FP32
transfer) with its energy spectrum: every pixel's null geodesic plus its
radiative transfer, the FP64 re-trace of the rays at the tau gate and the sum
over pixels, i.e. all of calculate_image's outputs (core.c:136-150).  Inputs
are resident in HBM before the timed region; outputs stay in HBM.

Prints one JSON line (rank 0).  Extra fields:
  roofline      the dominant kernel's (the persistent trace kernel's)
                algorithmic FLOP rate vs the FP64/FP32 vector peak for its mix
                of work, timed with HIP events around it on the library's
                stream; the render's and the approach kernel's beside it
  kernels       mean ms of each of the render's kernels
  grad          image + d/d(ln M_unit, R_high) throughput beside plain images
  cpu_baseline  the FP64 CPU oracle (OpenMP, all host cores) on the workload
  parity        GPU vs oracle on that same sample, every pixel
  configs       the other workloads: c2 (6 inclinations x 8 frequencies at
                512^2, one batched launch against 48 single-image launches),
                c3 (192x96x96 3D dump at 1024^2), c4 (image + gradient on the
                3D dump at 512^2), each with its ms, parity against the
                oracle on a pixel sample and roofline

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--cpu-sample N]
                    [--workload c1|c2|c3|c4] [--no-configs]
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
C2_INCL = [10.0, 25.0, 40.0, 55.0, 70.0, 85.0]
C2_FREQS = list(86e9 * (690.0 / 86.0) ** (np.arange(8) / 7.0))


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def _json(obj):
    """One JSON line; numpy scalars as Python numbers."""
    def scalar(o):
        if isinstance(o, np.generic):
            return o.item()
        raise TypeError("not JSON serializable: %r" % type(o))
    return json.dumps(obj, default=scalar)


def _dump(*shape):
    from raptor_amd import torus
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return torus.fishbone_moncrief(*shape)


def _parity(g, ora, gate):
    """GPU pixels g against the oracle's ora (same shape); gate: the oracle's
    distance of tau from the ln 1000 gate at its closest sample."""
    floor = 1e-6 * np.max(np.abs(ora))
    r = np.abs(g - ora) / np.maximum(np.abs(ora), floor)
    near = gate < 1e-4
    worst = np.argsort(r)[-3:][::-1]
    return {
        "pixels": int(g.shape[0]), "flux_rel": float(g.sum() / ora.sum() - 1.0),
        "max_pix_rel": float(r.max()), "n_pix_over_1e-3": int((r > 1e-3).sum()),
        "floor": "1e-6*max(I)", "worst": [[int(i), float(r[i])] for i in worst],
        "gate_adjacent": int(near.sum()),
        "max_pix_rel_gate_adjacent": float(r[near].max()) if near.any() else 0.0,
        "gate_note": "gate_adjacent: pixels whose FP64 tau passes within 1e-4 of the "
                     "ln 1000 gate at a sample; rays whose FP32 tau meets the gate where a "
                     "different decision would show are traced again with the FP64 transfer",
    }


def _sample(npix, n):
    stride = max(1, npix // n)
    return np.arange(0, npix, stride, dtype=np.int32)[:n], stride


def _roof(f64, f32, ms, name):
    from raptor_amd import flops
    peak = flops.mix_peak(f64, f32, FP64_PEAK_TFLOPS, FP32_PEAK_TFLOPS) if f64 + f32 else 0.
    achieved = (f64 + f32) / (ms * 1e-3) / 1e12 if ms > 0 else 0.
    return {"kernel": name, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak if peak else 0., "kernel_ms": ms,
            "alg_flops_fp64": f64, "alg_flops_fp32": f32}


def _timed(tr, fn, steps):
    """Mean kernel ms (HIP events), mean per-kernel ms when the render is one
    launch chunk, and wall ms per call of fn()."""
    kms, parts = [], []
    tr.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
        tr.synchronize()
        kms.append(tr.last_kernel_ms())
        try:
            parts.append(tr.last_kernel_times())
        except Exception:
            pass
    wall = 1e3 * (time.perf_counter() - t0) / steps
    per = {k: float(np.mean([p[k] for p in parts])) for k in parts[0]} if parts else None
    return float(np.mean(kms)), per, wall


def _counters(tr, incl, freqs, grad=False):
    n = len(incl) * tr.width * tr.height
    d_img = tr.alloc(n * len(freqs) * 8)
    d_dimg = tr.alloc(n * len(freqs) * 16) if grad else None
    d_cnt = tr.alloc(n * 8)
    try:
        tr.render_device(incl, freqs, d_img, d_dimg, d_cnt)
        tr.synchronize()
        return tr.d2h(d_cnt, (n, 2), np.int32)
    finally:
        tr.free(d_img)
        if d_dimg:
            tr.free(d_dimg)
        tr.free(d_cnt)


def _rooflines(cnt, nf, kms, per, grid_step=None, grad=False, kernel="trace kernel",
               lead=None):
    """Roofline of the render's kernels together and, given the rays' grid
    entry steps (the oracle's counters: gs for the sampled pixels pix, whose
    own step counts are st), of the trace kernel, the approach kernel and
    the ring order's lead kernel (lead: mask of the rays it traces, all their
    steps and samples) on their own.  With every ray sampled the approach /
    trace split of the other rays is exact; with a sample, the approach share
    of their steps is the sample's."""
    from raptor_amd import flops
    f64, f32 = flops.image_flops_split(cnt[:, 0], cnt[:, 1], nf, grad=grad)
    out = {"render": _roof(f64, f32, kms, "all kernels of the render"),
           "rk2_steps_per_launch": int(cnt[:, 0].sum()),
           "samples_per_launch": int(cnt[:, 1].sum())}
    if per is None or grid_step is None:
        return out
    gs, st, pix = grid_step
    main = np.ones(cnt.shape[0], bool) if lead is None else ~lead
    if lead is not None and lead.any():
        l64, l32 = flops.image_flops_split(cnt[lead, 0], cnt[lead, 1], nf, grad=grad)
        out["lead"] = _roof(l64, l32, per["lead"], "trace_kernel<1,...,false,0> (lead)")
        out["lead_rays"] = int(lead.sum())
    m64, m32 = flops.image_flops_split(cnt[main, 0], cnt[main, 1], nf, grad=grad)
    sm = main[pix]
    if pix.size == cnt.shape[0]:
        (a64, _), (t64, t32) = flops.kernel_split(cnt[main, 0], gs[sm], cnt[main, 1], nf,
                                                  grad=grad)
        out["split_from"] = "per-ray grid entry steps of the oracle (all rays)"
    else:
        share = float(np.where(gs[sm] >= 0, gs[sm], st[sm]).sum()) / float(st[sm].sum())
        a64 = share * m64
        t64, t32 = m64 - a64, m32
        out["split_from"] = ("approach share of the steps %.4f from the oracle's sample of "
                             "%d rays" % (share, int(sm.sum())))
    out["trace"] = _roof(t64, t32, per["trace"], kernel)
    out["approach"] = _roof(a64, 0.0, per["approach"], "approach_kernel<0>")
    return out


def _lead_mask(tr, incl, per):
    """Rays the ring order's lead kernel traced in the last render (None if it
    did not run)."""
    if not per or per.get("lead", 0.0) <= 0.0:
        return None
    order, nl = tr.claim_order(incl)
    m = np.zeros(order.size, bool)
    m[order[:nl]] = True
    return m


def bench_c1(args):
    from raptor_amd import flops, rt
    W = args.size
    dump = _dump(256, 256)
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
    d_spec = tr.alloc(8)

    # counters pass (untimed): per-ray steps and transfer samples
    cnt = _counters(tr, incl, freq)
    for _ in range(args.warmup):
        tr.render_device(incl, freq, d_img, d_spectrum=d_spec)
    kms, per, ms_per_step = _timed(
        tr, lambda: tr.render_device(incl, freq, d_img, d_spectrum=d_spec), args.steps)
    value = npix / (ms_per_step * 1e-3)

    # gradient leg: image + its two parameter derivatives (thermal fit only)
    grad = None
    if thermal and args.grad_steps > 0:
        tr.render_device(incl, freq, d_img, d_dimg, d_spectrum=d_spec)
        gk, gper, gwall = _timed(
            tr, lambda: tr.render_device(incl, freq, d_img, d_dimg, d_spectrum=d_spec),
            args.grad_steps)
        grad = {"image_plus_grad_per_s": 1e3 / gwall, "image_per_s": 1e3 / ms_per_step,
                "ratio": ms_per_step / gwall, "kernel_ms": gk, "kernels": gper,
                "achieved_tflops": flops.image_flops(cnt[:, 0], cnt[:, 1], 1, grad=True) /
                (gk * 1e-3) / 1e12}
        tr.render_device(incl, freq, d_img, d_spectrum=d_spec)   # the image parity reads
        tr.synchronize()
    img = tr.d2h(d_img, (npix,), np.float64)
    spec = float(tr.d2h(d_spec, (1,), np.float64)[0])
    jy = tr.jansky_factor()
    launch = tr.get_launch_config()
    split = launch["approach"] > 0 or (launch["approach"] == 0 and npix >= 1 << 20)
    lead = _lead_mask(tr, incl[0], per)
    trace_name = ("trace_kernel_3w<1,false,false,0,%s,0>" % ("true" if split else "false"))

    result = {
        "metric": METRIC, "value": value, "unit": "pixel-geodesics/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (seeded Fishbone-Moncrief torus, HARM MKS, a=0.9375)",
        "config": dict(WORKLOAD, image="%dx%d" % (W, W), emission=args.emission),
        "launch": dict(launch, approach_effective=100 if split else -1),
        "images_per_s": 1e3 / ms_per_step, "flux_jy": spec * jy,
        "spectrum_equals_image_sum": bool(abs(spec / img.sum() - 1) < 1e-12),
        "kernels": per, "kernel_ms": kms,
    }

    grid_step = None
    if args.cpu_sample > 0:
        from oracle import oracle as O
        n = min(args.cpu_sample, npix)
        pix, stride = _sample(npix, n)
        P = O.params(dump.grid, width=W, height=W, inclination=incl[0],
                     M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                     R_low=WORKLOAD["R_low"], cam_size=WORKLOAD["fov_M"],
                     stepsize=WORKLOAD["stepsize"], emission=emission)
        cores = host_cores()
        c0 = time.perf_counter()
        ora, ocnt, gate = O.render(P, dump.prims, freq, pixels=pix, counters=True, gate=True,
                                   nthreads=cores)
        c1 = time.perf_counter()
        cpu_rate = n / (c1 - c0)
        result["cpu_baseline"] = {
            "value": cpu_rate, "unit": "pixel-geodesics/s", "cores": cores, "kind": "port",
            "sample": ("all %d pixels" % npix if n == npix else
                       "%d of %d pixels (every %d-th)" % (n, npix, stride)) +
                      " of the same workload, FP64 oracle, OpenMP dynamic schedule, "
                      "%.1f s wall" % (c1 - c0)}
        result["speedup_vs_cpu"] = value / cpu_rate
        result["parity"] = _parity(img[pix], ora[:, 0], gate[:, 0])
        # a ray whose tau passed ln 1000 ends there (nothing later reaches its
        # pixel): its steps lie after the oracle's opaque step, within its trace
        st, op = cnt[pix, 0], ocnt[:, 5]
        agree = np.where(op >= 0, (st > op) & (st <= ocnt[:, 0]), st == ocnt[:, 0])
        result["parity"]["steps_agree_oracle_frac"] = float(np.mean(agree))
        result["parity"]["opaque_frac"] = float(np.mean(op >= 0))
        grid_step = (ocnt[:, 4], cnt[pix, 0], pix)

    roof = (_rooflines(cnt, 1, kms, per, grid_step, kernel=trace_name, lead=lead)
            if thermal else None)
    traffic = issue = None
    prof = {}
    tfile = os.path.join(ROOT, "profiles", "trace_kernel_traffic.json")
    if roof and W == 1024 and split and os.path.exists(tfile):
        with open(tfile) as fh:
            prof = json.load(fh)
        traffic = prof.get("trace_bytes_per_launch")
        insts = prof.get("trace_valu_wave_insts_per_launch")
        if insts and per:
            # a wave64 VALU instruction occupies its SIMD for 4 cycles
            issue = insts * 4 / (per["trace"] * 1e-3 * 1024 * 2.4e9)
    if roof and "trace" in roof:
        t = roof["trace"]
        result["roofline"] = {
            "bound": "valu", "achieved": t["achieved"], "peak": t["peak"], "unit": "TFLOP/s",
            "frac": t["frac"], "traffic": traffic,
            "traffic_note": "HBM bytes per launch of the trace kernel from rocprofv3 "
                            "FETCH_SIZE/WRITE_SIZE, copy-calibrated "
                            "(profiles/trace_kernel_traffic.json)",
            "traffic_algorithmic": prof.get("trace_algorithmic_bytes"),
            "render_traffic": prof.get("render_bytes"),
            "render_traffic_algorithmic": prof.get("render_algorithmic_bytes"),
            "kernel": t["kernel"], "kernel_ms": t["kernel_ms"],
            "kernel_note": "the dominant kernel alone, HIP events around it on the library "
                           "stream; algorithmic work = its rays' RK2 steps (from the "
                           "grid-entry step on) x 566 FP64 + samples x 300 FP32; the rays "
                           "near the critical curve or the spin axis are the lead kernel's "
                           "(second stream, beside the approach kernel)",
            "valu_issue_frac": issue,
            "peak_note": "FP64/FP32 mix-weighted vector peak (78.6 / 157.3 TFLOP/s)",
            "alg_flops_fp64": t["alg_flops_fp64"], "alg_flops_fp32": t["alg_flops_fp32"],
            "render": roof["render"], "approach": roof["approach"], "lead": roof.get("lead"),
            "lead_rays": roof.get("lead_rays"),
            "split_from": roof["split_from"],
            "rk2_steps_per_launch": roof["rk2_steps_per_launch"],
            "samples_per_launch": roof["samples_per_launch"],
        }
    elif roof:
        result["roofline"] = dict(roof["render"], bound="valu", traffic=None,
                                  kernel_note="the render's kernels together (no oracle "
                                              "counters to split them)")
    else:
        result["roofline"] = None
        result["roofline_note"] = ("no roofline: the frozen operation counts "
                                   "(raptor_amd/flops.py) are the thermal fit's")
    result["grad"] = grad
    tr.free(d_img)
    tr.free(d_dimg)
    tr.free(d_spec)
    tr.close()
    return result


def bench_c2(args):
    """Batched: 6 inclinations x 8 log-spaced frequencies (86-690 GHz) at 512^2
    in one persistent launch (autotuned shape) against the same 48 images as
    single-image launches replayed from one HIP graph."""
    from raptor_amd import autotune, rt
    dump = _dump(256, 256)
    W = 512
    incl, freqs = C2_INCL, C2_FREQS
    tr = rt.Tracer(0)
    tr.set_params(width=W, height=W, M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                  R_low=WORKLOAD["R_low"])
    tr.load_dump(dump)
    npix = W * W
    tune = autotune.tune(tr, incl, freqs, reps=1, path=None)
    nimg = len(incl) * len(freqs)
    d_b = tr.alloc(len(incl) * npix * len(freqs) * 8)
    d_s = tr.alloc(nimg * npix * 8)
    d_spec = tr.alloc(nimg * 8)
    for _ in range(max(1, args.warmup)):
        tr.render_device(incl, freqs, d_b, d_spectrum=d_spec)
    kb, perb, wb = _timed(tr, lambda: tr.render_device(incl, freqs, d_b, d_spectrum=d_spec),
                          args.steps)
    # 48 single-image launches, captured once and replayed
    tune_cfg = tr.get_launch_config()
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
    ks, _, ws = _timed(tr, g.launch, args.steps)
    b = tr.d2h(d_b, (len(incl), npix, len(freqs)), np.float64)
    s = tr.d2h(d_s, (len(incl), len(freqs), npix), np.float64)
    spec = tr.d2h(d_spec, (len(incl), len(freqs)), np.float64)
    same = bool(np.array_equal(b, np.transpose(s, (0, 2, 1))))
    tr.set_launch_config(**{k_: v for k_, v in tune_cfg.items() if k_ != "block"})
    cnt = _counters(tr, incl, freqs)
    g.destroy()
    tr.free(d_b)
    tr.free(d_s)
    tr.free(d_spec)
    tr.close()
    rays = len(incl) * npix
    out = {"workload": "c2", "unit": "pixel-geodesics/s", "value": rays / (wb * 1e-3),
           "ms_per_step": wb, "steps": args.steps,
           "config": dict(workload="harm2d-torus-256x256_img512x512_6incl_8freq",
                          dump="256x256x1", image="512x512", inclinations=incl,
                          freqs_hz=freqs),
           "batched": {"kernel_ms": kb, "kernels": perb, "wall_ms": wb,
                       "launch_config": tune["config"],
                       "images_per_s": nimg / (wb * 1e-3),
                       "roofline": _rooflines(cnt, len(freqs), kb, None)["render"]},
           "per_image_graph": {"launches": nimg, "kernel_ms": ks, "wall_ms": ws,
                               "images_per_s": nimg / (ws * 1e-3)},
           "speedup_batched": ws / wb, "batched_equals_per_image": same,
           "spectrum_equals_image_sum": bool(np.allclose(spec, b.sum(axis=1), rtol=1e-12,
                                                         atol=0))}
    if args.cpu_sample > 0:
        # the FP64 oracle on every 64th pixel of each inclination, all 8
        # frequencies: the CPU baseline, and parity for all 48 images
        from oracle import oracle as O
        pix, stride = _sample(npix, npix // 64)
        cores = host_cores()
        oras, gates, ocnts = [], [], []
        c0 = time.perf_counter()
        for inc in incl:
            P = O.params(dump.grid, width=W, height=W, inclination=inc,
                         M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                         R_low=WORKLOAD["R_low"])
            o, oc, gt = O.render(P, dump.prims, freqs, pixels=pix, counters=True, gate=True,
                                 nthreads=cores)
            oras.append(o)
            gates.append(gt)
            ocnts.append(oc)
        c1 = time.perf_counter()
        n = len(incl) * pix.size
        out["cpu_baseline"] = {
            "value": n / (c1 - c0), "unit": "pixel-geodesics/s", "cores": cores,
            "kind": "port",
            "sample": "every %d-th pixel of each of the %d inclinations (%d rays, %d "
                      "frequencies each), FP64 oracle, OpenMP dynamic schedule, %.1f s "
                      "wall" % (stride, len(incl), n, len(freqs), c1 - c0)}
        par = _parity(b[:, pix, :].reshape(-1), np.stack(oras).reshape(-1),
                      np.stack(gates).reshape(-1))
        par["images"] = nimg
        par["max_image_flux_rel"] = float(np.max(np.abs(
            b[:, pix, :].sum(axis=1) / np.stack(oras).sum(axis=1) - 1.0)))
        out["parity"] = par
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def bench_3d(args, which):
    """c3: the 3D 192x96x96 torus (m=1..4 azimuthal modes, 10%) at 1024^2, 230
    GHz.  c4: the same dump at 512^2, image + d/d(ln M_unit, R_high) against
    the image alone, with a central-difference check of the gradient."""
    from raptor_amd import rt
    dump = _dump(192, 96, 96)
    W = 1024 if which == "c3" else 512
    npix = W * W
    incl, freq = [WORKLOAD["inclination_deg"]], [WORKLOAD["freq_hz"]]
    tr = rt.Tracer(0)
    tr.set_params(width=W, height=W, M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                  R_low=WORKLOAD["R_low"])
    tr.load_dump(dump)
    d_img = tr.alloc(npix * 8)
    d_dimg = tr.alloc(npix * 16)
    d_spec = tr.alloc(8)
    for _ in range(max(1, args.warmup)):
        tr.render_device(incl, freq, d_img, d_spectrum=d_spec)
    ki, peri, wi = _timed(tr, lambda: tr.render_device(incl, freq, d_img, d_spectrum=d_spec),
                          args.steps)
    cnt = _counters(tr, incl, freq)
    out = {"workload": which, "unit": "pixel-geodesics/s", "value": npix / (wi * 1e-3),
           "ms_per_step": wi, "steps": args.steps,
           "config": dict(workload="harm3d-torus-192x96x96_img%dx%d_230GHz_i60" % (W, W),
                          dump="192x96x96", image="%dx%d" % (W, W),
                          prims_MB=dump.prims.size * 4 / 1e6),
           "image": {"kernel_ms": ki, "kernels": peri, "wall_ms": wi,
                     "images_per_s": 1e3 / wi}}
    img = tr.d2h(d_img, (npix,), np.float64)
    lead = _lead_mask(tr, incl[0], peri)
    if which == "c4":
        tr.render_device(incl, freq, d_img, d_dimg, d_spectrum=d_spec)
        kg, perg, wg = _timed(
            tr, lambda: tr.render_device(incl, freq, d_img, d_dimg, d_spectrum=d_spec),
            args.steps)
        an = tr.d2h(d_dimg, (npix, 2), np.float64)
        gcnt = _counters(tr, incl, freq, grad=True)
        out["grad"] = {"kernel_ms": kg, "kernels": perg, "wall_ms": wg,
                       "image_plus_grad_per_s": 1e3 / wg, "ratio_vs_image": wi / wg}
    tr.free(d_img)
    tr.free(d_dimg)
    tr.free(d_spec)
    tr.close()
    grid_step = None
    if args.cpu_sample > 0:
        # the FP64 oracle on every 64th pixel: the CPU baseline and parity
        from oracle import oracle as O
        pix, stride = _sample(npix, npix // 64)
        P = O.params(dump.grid, width=W, height=W, inclination=incl[0],
                     M_unit=WORKLOAD["M_unit"], R_high=WORKLOAD["R_high"],
                     R_low=WORKLOAD["R_low"])
        cores = host_cores()
        c0 = time.perf_counter()
        ora0, ocnt, gate0 = O.render(P, dump.prims, freq, pixels=pix, counters=True,
                                     gate=True, nthreads=cores)
        c1 = time.perf_counter()
        out["cpu_baseline"] = {
            "value": pix.size / (c1 - c0), "unit": "pixel-geodesics/s", "cores": cores,
            "kind": "port",
            "sample": "every %d-th pixel (%d rays), FP64 oracle, OpenMP dynamic schedule, "
                      "%.1f s wall" % (stride, pix.size, c1 - c0)}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["parity"] = _parity(img[pix], ora0[:, 0], gate0[:, 0])
        grid_step = (ocnt[:, 4], cnt[pix, 0], pix)
    split = npix >= 1 << 20
    out["image"]["roofline"] = _rooflines(
        cnt, 1, ki, peri if split else None, grid_step if split else None,
        kernel="trace_kernel_3w<1,false,true,0,true,0>", lead=lead)
    if which == "c4":
        out["grad"]["roofline"] = _rooflines(gcnt, 1, kg, None, grad=True)["render"]
    if which == "c4" and args.cpu_sample > 0:
        # the gradient against central differences (step 1e-3 and 5e-4) of the
        # FP64 oracle on 2048 sampled pixels; GPU-image differences would be
        # FP32-noise limited
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
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grad-steps", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20,
                    help="pixels of the headline image traced by the CPU oracle (default: "
                         "all of the 1024^2 image, about 10 s on 16 cores; 0 = skip); the "
                         "other workloads trace every 64th pixel")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--emission", choices=["thsynch", "thsynchav", "kappa"], default="thsynch",
                    help="emission model of the c1 image (default: thsynch, the path's); "
                         "the others have no gradient leg and no roofline (the frozen "
                         "operation counts are the thermal fit's)")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4"], default="c1",
                    help="c1: the headline 1024^2 image (default, with the other workloads "
                         "under 'configs'); c2: 6 inclinations x 8 frequencies at 512^2, one "
                         "launch vs 48; c3: 3D 192x96x96 dump at 1024^2; c4: image+gradient "
                         "on the 3D dump at 512^2")
    ap.add_argument("--config-steps", type=int, default=5,
                    help="timed steps of each of the other workloads in the default run")
    ap.add_argument("--no-configs", action="store_true",
                    help="the headline workload only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(pixels, cameras and frequencies shard without communication). "
              "Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c2":
        print(_json(bench_c2(args)), flush=True)
        return
    if args.workload in ("c3", "c4"):
        print(_json(bench_3d(args, args.workload)), flush=True)
        return
    result = bench_c1(args)
    if not args.no_configs and args.size == 1024 and args.emission == "thsynch":
        sub = argparse.Namespace(**vars(args))
        sub.steps = args.config_steps
        sub.warmup = 1
        result["configs"] = {"c2": bench_c2(sub), "c3": bench_3d(sub, "c3"),
                             "c4": bench_3d(sub, "c4")}
        result["configs_note"] = ("each workload's own ms, parity against the FP64 oracle on "
                                  "every 64th pixel and roofline; not part of value")
    print(_json(result), flush=True)


if __name__ == "__main__":
    main()

"""Autotuner for the persistent trace kernel's launch shape.

Knobs (rt_launch_config): resident workgroups per CU (occupancy), the wave
refill threshold, the T x T pixel tiles in which rays are handed out, the
order of the tiles (centre-out or row-major) and the share of rays whose
camera-to-grid leg runs in the approach kernel (all or none).  `tune` times
every candidate on the caller's own workload (same dump,
image size, cameras, frequencies, gradient flag) with the library's HIP-event
kernel timer, keeps the fastest, applies it to the tracer and stores it in a
JSON cache keyed by the workload class, so later runs start from it
(`apply_cached`).
"""
from __future__ import annotations

import itertools
import json
import os

import numpy as np

from .rt import APPROACH_NONE, ORDER_RING, ORDER_SCAN

CACHE = os.environ.get("RAPTOR_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "raptor_amd",
                                    "autotune.json"))

DEFAULT_SPACE = dict(blocks_per_cu=(0, 4, 6), refill=(1, 4, 32), tile=(1, 4, 8, 16),
                     order=(ORDER_RING, ORDER_SCAN), approach=(APPROACH_NONE, 100))


def workload_key(tracer, n_cam, n_freq, grad, dim3):
    return "%dx%d:cam%d:nf%d:%s:%s" % (tracer.width, tracer.height, n_cam, n_freq,
                                      "grad" if grad else "image", "3d" if dim3 else "2d")


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(path, key, entry):
    data = _load(path)
    data[key] = entry
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1)
    os.replace(tmp, path)


def apply_cached(tracer, n_cam, n_freq, grad=False, dim3=False, path=CACHE):
    """Apply a cached best configuration; returns it or None."""
    entry = _load(path).get(workload_key(tracer, n_cam, n_freq, grad, dim3))
    if not entry:
        return None
    tracer.set_launch_config(**entry["config"])
    return entry["config"]


def _time(tracer, cfg, incl, fr, d_img, d_dimg, reps):
    """Median kernel ms of `reps` renders with launch configuration cfg (after
    one untimed render that builds whatever the shape needs)."""
    tracer.set_launch_config(**cfg)
    tracer.render_device(incl, fr, d_img, d_dimg)
    tracer.synchronize()
    ms = []
    for _ in range(reps):
        tracer.render_device(incl, fr, d_img, d_dimg)
        tracer.synchronize()
        ms.append(tracer.last_kernel_ms())
    return float(np.median(ms))


def tune(tracer, inclinations, freqs, grad=False, dim3=False, space=None, reps=2,
         path=CACHE, refine=4, refine_reps=5):
    """Time every configuration of `space` on this workload, then the `refine`
    fastest again with `refine_reps` renders each (one sweep's minimum over
    many shapes is biased towards a lucky timing: differences between the
    leading shapes are a few percent, like the run-to-run noise); apply and
    cache the fastest of those.  Returns {"config", "ms", "table",
    "refined"}."""
    space = dict(DEFAULT_SPACE, **(space or {}))
    incl = np.atleast_1d(inclinations)
    fr = np.atleast_1d(freqs)
    n = incl.size * tracer.width * tracer.height * fr.size
    d_img = tracer.alloc(n * 8)
    d_dimg = tracer.alloc(2 * n * 8) if grad else None
    table = []
    try:
        keys = sorted(space)
        for values in itertools.product(*(space[k] for k in keys)):
            cfg = dict(zip(keys, values))
            table.append((cfg, _time(tracer, cfg, incl, fr, d_img, d_dimg, reps)))
        lead = sorted(table, key=lambda t: t[1])[:max(1, refine)]
        refined = [(cfg, _time(tracer, cfg, incl, fr, d_img, d_dimg, refine_reps))
                   for cfg, _ in lead]
    finally:
        tracer.free(d_img)
        if d_dimg:
            tracer.free(d_dimg)
    best_cfg, best_ms = min(refined, key=lambda t: t[1])
    tracer.set_launch_config(**best_cfg)
    if path:
        _store(path, workload_key(tracer, incl.size, fr.size, grad, dim3),
               {"config": best_cfg, "ms": best_ms})
    return {"config": best_cfg, "ms": best_ms, "table": table, "refined": refined}

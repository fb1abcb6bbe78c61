"""Algorithmic work of the per-pixel path, for the roofline of the trace kernel.

FLOPs of one image = sum over rays of
    N_step * F_STEP + N_rt * (F_FLUID + n_freq * F_NU)   (+ grad terms)
where N_step (RK2 steps) and N_rt (samples that pass the grid gate and the
sigma cut) are counted per ray on the device.

The constants are frozen hand counts of the reference's algorithm in its
minimal form: the MKS connection (metric.c:538-638) evaluated from its
unique entries and contracted with k^a k^b using its symmetry, the
trigonometry of a point computed once, no repeated metric evaluations
(raptor_harm_model.c and radiative_transfer.c each evaluate the metric
twice per sample; counted once).  Each add, subtract or multiply counts 1
(an FMA 2), and each division, square/cube root, exp, sin or cos counts 1;
products of literal constants are folded.  Per-function breakdown:

  make_point            5 arith + 5 transcendental (exp, 2 sincospi) = 10
  geodesic_accel      249 arith + 1 division                         = 250
  stepsize              3 arith + 2 division                         = 5
  rk2_step            2*accel + 2*make_point + 40 updates            = 560
  loop bookkeeping      1                                             = 1
  F_STEP = 5 + 560 + 1                                                = 566
    of which 10 transcendental, 4 division

  fluid_sample (FP32) 231 arith + 9 reciprocal + 2 sqrt              = 242 (F_FLUID)
  transfer per freq    46 arith + 4 reciprocal + 1 sqrt + 7 transc.  = 58  (F_NU)
  tangent pass: +11 per sample (dTheta_e/dR_high), +41 per sample-frequency

The counts stay fixed when the implementation does the same mathematics
with less arithmetic, so such changes raise the achieved rate: the tracer
forms the contraction from the Kerr-Schild structure of the metric (about
120 operations instead of 250), derives the trigonometry of the RK2
midpoint and end point by angle addition from the start point, and replaces
FP64 divisions by a refined reciprocal.
"""

F_STEP = 566
F_STEP_TRANS = 10
F_STEP_DIV = 4
F_FLUID = 242
F_NU = 58
F_FLUID_GRAD = 11
F_NU_GRAD = 41


def image_flops_split(steps, samples, n_freq, grad=False):
    """(FP64 FLOPs, FP32 FLOPs): the geodesic steps run in FP64, the fluid
    sample and transfer in FP32 (their FP64 tangent accumulators are small
    and counted with the FP32 part)."""
    import numpy as np
    ns = float(np.sum(steps))
    nr = float(np.sum(samples))
    f64 = ns * F_STEP
    f32 = nr * (F_FLUID + n_freq * F_NU)
    if grad:
        f32 += nr * (F_FLUID_GRAD + n_freq * F_NU_GRAD)
    return f64, f32


def kernel_split(steps, grid_step, samples, n_freq, grad=False):
    """Algorithmic work of the split launch's two kernels, given per-ray
    steps, grid_step (steps taken before the first one that lands inside the
    grid, -1 if none does) and samples: ((FP64, FP32) of the approach
    kernel, (FP64, FP32) of the trace kernel).  The approach kernel takes a
    ray up to its grid entry (a ray that never reaches the grid, all its
    steps); the trace kernel repeats the entering step and takes every step
    and transfer sample after it.  The approach kernel's evaluation of the
    entering step, which the trace kernel repeats, is not algorithmic work."""
    import numpy as np
    steps = np.asarray(steps, dtype=np.float64)
    gs = np.asarray(grid_step, dtype=np.float64)
    enter = gs >= 0
    ns_app = float(np.sum(np.where(enter, gs, steps)))
    ns_tr = float(np.sum(np.where(enter, steps - gs, 0.0)))
    _, f32 = image_flops_split(0, samples, n_freq, grad=grad)
    return (ns_app * F_STEP, 0.0), (ns_tr * F_STEP, f32)


def mix_peak(f64, f32, peak64, peak32):
    """Peak rate for a mix of FP64 and FP32 work: total work over the time it
    takes when each part runs at its own peak."""
    return (f64 + f32) / (f64 / peak64 + f32 / peak32)


def image_flops(steps, samples, n_freq, grad=False):
    """Algorithmic FLOPs given per-ray step and sample counts (arrays or sums)."""
    import numpy as np
    ns = float(np.sum(steps))
    nr = float(np.sum(samples))
    f = ns * F_STEP + nr * (F_FLUID + n_freq * F_NU)
    if grad:
        f += nr * (F_FLUID_GRAD + n_freq * F_NU_GRAD)
    return f

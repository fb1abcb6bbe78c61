"""Algorithmic work of the per-pixel path, for the roofline of the trace kernel.

FLOPs of one image = sum over rays of
    N_step * F_STEP + N_rt * (F_FLUID + n_freq * F_NU)   (+ grad terms)
where N_step (RK2 steps) and N_rt (samples that pass the grid gate and the
sigma cut) are counted per ray on the device.

The constants are hand counts of the arithmetic in raptor_amd/csrc/rt_physics.hpp
(the minimal form of the reference's algorithm: the connection contracted from
its unique entries, trigonometry shared per point, no repeated metric
evaluations).  Each add, subtract or multiply counts 1 (an FMA 2), and each
division, square/cube root, exp, sin or cos counts 1; products of literal
constants are folded.  Per-function breakdown:

  make_point            6 arith + 5 transcendental            = 11
  geodesic_accel      235 arith + 5 division                  = 240
  stepsize             10 arith + 7 division                  = 17
  rk2_step            2*accel + 2*make_point + 40 updates     = 542
  loop bookkeeping      1                                      = 1
  F_STEP = 17 + 542 + 1                                        = 560
    of which 10 transcendental, 17 division

  fluid_sample        226 arith + 20 division + 2 sqrt         = 248 (F_FLUID)
  transfer per freq    40 arith + 10 division + 7 transc.      = 57  (F_NU)
  tangent pass: +15 per sample (dTheta_e/dR_high), +41 per sample-frequency
"""

F_STEP = 560
F_STEP_TRANS = 10
F_STEP_DIV = 17
F_FLUID = 248
F_NU = 57
F_FLUID_GRAD = 15
F_NU_GRAD = 41


def image_flops(steps, samples, n_freq, grad=False):
    """Algorithmic FLOPs given per-ray step and sample counts (arrays or sums)."""
    import numpy as np
    ns = float(np.sum(steps))
    nr = float(np.sum(samples))
    f = ns * F_STEP + nr * (F_FLUID + n_freq * F_NU)
    if grad:
        f += nr * (F_FLUID_GRAD + n_freq * F_NU_GRAD)
    return f

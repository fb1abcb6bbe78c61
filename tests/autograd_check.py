"""Helper for test_gpu_autograd.py, run in its own process so that torch is
imported before the library (see raptor_amd/autograd.py).  Prints one JSON
line: the autograd gradient of a weighted image sum and central differences
(step 1e-3) of the same sum from the FP64 oracle.  Pixels whose oracle
difference quotient changes between steps h and h/2 (a gate flips inside the
step, making the image locally discontinuous) get weight 0."""
import json
import math
import os
import sys
import warnings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from raptor_amd import autograd  # noqa: E402  (imports torch first)
import numpy as np  # noqa: E402
import torch  # noqa: E402

from oracle import oracle as O  # noqa: E402
from raptor_amd import rt, torus  # noqa: E402


def main():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        dump = torus.fishbone_moncrief(48, 40, 12)
    W, incl, freq, M0, R0 = 20, 60.0, 2.3e11, 1e23, 10.0

    def oracle_img(M, R):
        P = O.params(dump.grid, width=W, height=W, inclination=incl, M_unit=M, R_high=R)
        return O.render(P, dump.prims, [freq])[:, 0]

    h = 1e-3
    steps = lambda q, hh: ((M0 * math.exp(hh), R0), (M0 * math.exp(-hh), R0), 2 * hh) if q == 0 \
        else ((M0, R0 * (1 + hh)), (M0, R0 * (1 - hh)), 2 * hh * R0)
    fdpix = []
    smooth = np.ones(W * W, bool)
    for q in range(2):
        quot = []
        for hh in (h, h / 2):
            a, b, d = steps(q, hh)
            quot.append((oracle_img(*a) - oracle_img(*b)) / d)
        den = np.maximum(np.abs(quot[0]), 1e-3 * np.abs(quot[0]).max())
        smooth &= np.abs(quot[0] - quot[1]) / den < 1e-3
        fdpix.append(quot[0])
    wn = np.random.default_rng(3).uniform(0.5, 1.5, W * W) * smooth
    w = torch.from_numpy(wn.reshape(1, W * W, 1))
    with rt.Tracer(0) as tr:
        # the tracer's own parameters differ from the ones the render is asked
        # for (and use another integrator): the render takes its own and
        # keeps the integrator, and the tracer's come back afterwards
        tr.set_params(width=W, height=W, M_unit=3 * M0, R_high=2 * R0, integrator=rt.INT_VERLET)
        tr.load_dump(dump)
        params = torch.tensor([math.log(M0), R0], dtype=torch.float64, requires_grad=True)
        img_v = autograd.render_image(params, tr, [incl], [freq]).detach()
        restored = (tr.params.M_unit == 3 * M0 and tr.params.R_high == 2 * R0 and
                    tr.params.integrator == rt.INT_VERLET)
        tr.set_params(width=W, height=W, M_unit=M0, R_high=R0)
        img = autograd.render_image(params, tr, [incl], [freq])
        verlet_differs = bool((img_v != img.detach()).any())
        loss = (img * w).sum()
        loss.backward()
        grad = params.grad.tolist()

    fd = [float((fdpix[q] * wn).sum()) for q in range(2)]
    print(json.dumps({"loss": float(loss), "loss_oracle": float((oracle_img(M0, R0) * wn).sum()),
                      "grad": grad, "fd": fd, "smooth_frac": float(smooth.mean()),
                      "params_restored": restored, "integrator_kept": verlet_differs}))


if __name__ == "__main__":
    main()

"""torch.autograd.Function over the library's image + gradient render, for
checking gradients (gradcheck-style) from PyTorch.

The forward renders the image with the tangent pass on; the backward
contracts the upstream gradient with the per-pixel Jacobian
d(image)/d(ln M_unit, R_high).  Tensors live on the CPU; the work runs on the
GPU through the C-ABI.

torch must be imported before libraptor_rt.so is loaded in a process: torch
ships its own HIP runtime with the same soname, and loading the library first
would put two HIP runtimes in one process.  Importing this module enforces
that order.
"""
from __future__ import annotations

import torch  # noqa: E402  (before raptor_amd.rt loads the library)

from . import rt

if rt._LIB is not None:
    raise ImportError("import raptor_amd.autograd (torch) before the library is loaded")


class RenderImage(torch.autograd.Function):
    """image = RenderImage.apply(params, tracer, inclinations, freqs)

    params: float64 tensor [ln M_unit, R_high].  Returns the image
    [n_cam, W*H, n_freq] (float64, I_nu in cgs)."""

    @staticmethod
    def forward(ctx, params, tracer, inclinations, freqs):
        p = params.detach().cpu().double()
        base = tracer.params
        keep = dict(width=base.width, height=base.height, R_low=base.R_low, MBH=base.MBH,
                    cam_size=base.cam_size_x, cam_size_y=base.cam_size_y,
                    stepsize=base.stepsize, max_steps=base.max_steps,
                    integrator=base.integrator, emission=base.emission)
        tracer.set_params(M_unit=float(torch.exp(p[0])), R_high=float(p[1]), **keep)
        try:
            img, dimg, _ = tracer.render(inclinations, freqs, grad=True)
        finally:
            # the tracer's own parameters again, whatever the render did
            tracer.set_params(M_unit=base.M_unit, R_high=base.R_high, **keep)
        ctx.save_for_backward(torch.from_numpy(dimg))
        return torch.from_numpy(img)

    @staticmethod
    def backward(ctx, grad_out):
        (dimg,) = ctx.saved_tensors
        g = (grad_out.double().unsqueeze(-1) * dimg).sum(dim=(0, 1, 2))
        return g, None, None, None


def render_image(params, tracer, inclinations, freqs):
    return RenderImage.apply(params, tracer, inclinations, freqs)

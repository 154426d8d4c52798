"""Differentiable rendering with respect to material colours.

`render(colors, tracer, cam, settings)` renders with the given per-material
colours (albedo; a light's colour is its emission, as in the reference) and
is differentiable through `rt_render_adjoint`: the backward pass replays the
same Philox paths (common random numbers) and scatters dL/dI into d colour.
Everything stays on the device; the C ABI is called with torch pointers.
"""
from __future__ import annotations

import torch

from .tracer import Tracer


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, colors: torch.Tensor, tracer: Tracer, cam, settings):
        assert colors.is_cuda and colors.dtype == torch.float32 and colors.dim() == 2 and colors.shape[1] == 3
        colors = colors.contiguous()
        tracer.set_stream(torch.cuda.current_stream(colors.device).cuda_stream)
        tracer.set_material_colors(colors)
        img = torch.empty((cam.height, cam.width, 3), dtype=torch.float32, device=colors.device)
        tracer.render(cam, settings, rgb=img)
        ctx.tracer, ctx.cam, ctx.settings = tracer, cam, settings
        ctx.save_for_backward(colors)
        return img

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        tracer = ctx.tracer
        (colors,) = ctx.saved_tensors
        # the context may have rendered other colours since the forward pass
        tracer.set_stream(torch.cuda.current_stream(grad_img.device).cuda_stream)
        tracer.set_material_colors(colors)
        d_color = torch.zeros_like(colors)
        tracer.render_adjoint(ctx.cam, ctx.settings, grad_img.contiguous(), d_color)
        return d_color, None, None, None


def render(colors: torch.Tensor, tracer: Tracer, cam, settings) -> torch.Tensor:
    return RenderFunction.apply(colors, tracer, cam, settings)

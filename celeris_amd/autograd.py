"""torch.autograd bridge: the forward sum and its hand-written adjoint kernel.

``rs_field(t, px, py, planes)`` returns the complex field E (complex128, all planes
concatenated) for pillar transmissions ``t`` (complex128 [nsets, npil]).  Its
backward is the adjoint kernel: for a real loss L, torch hands backward
grad_E = dL/dRe(E) + i dL/dIm(E), and the vector-Jacobian product of the linear
map E = G t is G^H grad_E, which is exactly crs_adjoint.  Gradients with respect
to pillar phase and amplitude follow by torch's chain rule through
``t = a * exp(i phi)``.  Positions are not differentiated.
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from .rs import Context, Plane

_contexts = {}


def _context(device: int) -> Context:
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    return ctx


class RSField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t, px, py, planes, mode, context):
        if t.dim() == 1:
            t2 = t.reshape(1, -1)
        else:
            t2 = t
        field, _ = context.forward(px, py, t2.contiguous(), planes, mode=mode, field=True)
        ctx.save_for_backward(px, py)
        ctx.planes = planes
        ctx.mode = mode
        ctx.context = context
        ctx.t_shape = t.shape
        ctx.nsets = t2.shape[0]
        return field

    @staticmethod
    def backward(ctx, grad_field):
        px, py = ctx.saved_tensors
        g = ctx.context.adjoint(px, py, grad_field.contiguous(), ctx.planes, nsets=ctx.nsets,
                                mode=ctx.mode)
        return g.reshape(ctx.t_shape), None, None, None, None, None


def rs_field(t: torch.Tensor, px: torch.Tensor, py: torch.Tensor, planes: Sequence[Plane],
             mode: str = "fp32c", context: Optional[Context] = None) -> torch.Tensor:
    """Differentiable complex field; see the module docstring."""
    if context is None:
        context = _context(px.device.index or 0)
    return RSField.apply(t, px, py, list(planes), mode, context)


def encircled_energy(intensity: torch.Tensor, plane: Plane, radius: float,
                     cx: float = 0.0, cy: float = 0.0) -> torch.Tensor:
    """Fraction of the window's power within `radius` of (cx, cy): the reference's
    encircled-energy ratio in/total (focal.cpp:126-137), differentiable."""
    n_y, n_x = plane.ny, plane.nx
    i = torch.arange(n_x, dtype=torch.float64, device=intensity.device)
    j = torch.arange(n_y, dtype=torch.float64, device=intensity.device)
    fx = plane.x0 + i * plane.sx
    fy = plane.y0 + j * plane.sy
    r = torch.sqrt((fx[None, :] - cx) ** 2 + (fy[:, None] - cy) ** 2)
    img = intensity.reshape(n_y, n_x)
    inside = (r <= radius).to(img.dtype)
    return (img * inside).sum() / img.sum()

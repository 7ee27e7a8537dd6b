"""Differentiable regridding for PyTorch: the counterpart of a ChainRulesCore
``rrule`` for ``regrid!`` (the reference has none).

``regrid_ad(dst_points, src_points, s)`` builds the regridder from the two
point grids (``(ni+1, nj+1, 3)`` or ``(nblocks, ni+1, nj+1, 3)`` unit vectors,
any device), applies it to ``s`` (FP64, cells along dim 0) and back-propagates
into ``s`` and into both point arrays.  The forward is the device build + FP64
apply; the backward is the transpose apply for ``s`` and
``cr_regrid_backward``'s intersection-area derivatives for the points.
"""
from __future__ import annotations

import torch

from .grids import Grid
from .regridder import Regridder, regrid, regrid_backward


class RegridFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, dst_points, src_points, threshold_rel=1e-14):
        dst = Grid.from_points(dst_points.detach().cpu().numpy())
        src = Grid.from_points(src_points.detach().cpu().numpy())
        R = Regridder(dst, src, radius=1.0, threshold_rel=threshold_rel)
        y = regrid(R, s.detach())
        ctx.saved = (R, dst, src, s.detach(), y)
        ctx.shapes = (dst_points.shape, src_points.shape)
        ctx.devices = (dst_points.device, src_points.device)
        return y

    @staticmethod
    def backward(ctx, gy):
        R, dst, src, s, y = ctx.saved
        need_pts = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        gx, g_dst, g_src = regrid_backward(R, dst, src, s, y, gy.contiguous(), field_grad=ctx.needs_input_grad[0],
                                           vertex_grad=need_pts)
        gd = gs = None
        if need_pts:
            gd = _to_points(g_dst, ctx.shapes[0]).to(ctx.devices[0])
            gs = _to_points(g_src, ctx.shapes[1]).to(ctx.devices[1])
        return gx, gd, gs, None


def _to_points(g, shape):
    """Caller-vertex gradients (point blocks, i fastest) -> the points' shape."""
    if len(shape) == 3:  # (ni+1, nj+1, 3)
        W, H = shape[0], shape[1]
        return g.view(H, W, 3).permute(1, 0, 2)
    nb, W, H = shape[0], shape[1], shape[2]
    return g.view(nb, H, W, 3).permute(0, 2, 1, 3)


def regrid_ad(dst_points, src_points, s, threshold_rel: float = 1e-14):
    """Differentiable y = regrid(Regridder(dst, src), s) on the unit sphere."""
    return RegridFunction.apply(s, dst_points, src_points, threshold_rel)

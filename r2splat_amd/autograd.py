"""Differentiable ``rasterize`` / ``voxelize`` on top of the explicit calls.

``torch.autograd.Function`` wrappers whose backward is the hand-written VJP
(``explicit.rasterize_bwd`` / ``voxelize_bwd``): the forward returns the image
(or volume), or ``(mean, variance)`` with ``uncertainty``, and gradients flow
to ``means``, ``quats``, ``scales``, ``densities`` and, when given, ``beta``
and ``uncertainty``. Cameras and the volume frame are constants.

``recompute`` chooses between holding the per-intersection lists from the
forward to the backward (False) and rebuilding them in the backward (True).
``None`` decides per call: the lists are rebuilt when their estimated size
exceeds ``recompute_budget`` bytes (default: a quarter of the device's
memory), which is the regime of many views times many primitives.

``mse_loss`` is the mean squared error of a float32 image against a target
in one kernel that also writes its adjoint 2 (image - target) / n (plus a
fixed-order reduction of per-block partial sums), so an MSE step adds two
launches and an elementwise scale in the backward.

The reference exposes only the explicit (output, ctx) pair
(r2splat/explicit.py:102-351); these wrappers compose it with autograd.
"""

from __future__ import annotations

from typing import Optional, Sequence

import torch

from . import explicit

Tensor = torch.Tensor

# Bytes held per intersection between forward and backward (record + keys),
# and per (camera, primitive) (projected state), used by the recompute policy.
_BYTES_PER_ISECT = 48
_BYTES_PER_PROJECTION = 96
_TILES_PER_PROJECTION = 8


def _estimate_bytes(n_prims: int, n_views: int) -> int:
    return n_views * n_prims * (_BYTES_PER_PROJECTION + _BYTES_PER_ISECT * _TILES_PER_PROJECTION)


def _should_recompute(recompute, budget, n_prims, n_views, device) -> bool:
    if recompute is not None:
        return bool(recompute)
    if budget is None:
        budget = torch.cuda.get_device_properties(device).total_memory // 4 if device.type == "cuda" else 0
    return _estimate_bytes(n_prims, n_views) > budget


class _Rasterize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, options, means, quats, scales, densities, beta, uncertainty):
        output, ectx = explicit.rasterize_fwd(means.detach(), quats.detach(), scales.detach(), densities.detach(),
                                              beta=None if beta is None else beta.detach(),
                                              uncertainty=None if uncertainty is None else uncertainty.detach(),
                                              **options)
        ctx.explicit = ectx
        ctx.has = (beta is not None, uncertainty is not None)
        return output

    @staticmethod
    def backward(ctx, *grad_outputs):
        v_image = grad_outputs[0]
        v_variance = grad_outputs[1] if len(grad_outputs) > 1 else None
        ectx, ctx.explicit = ctx.explicit, None
        if ectx is None:
            raise RuntimeError("rasterize: backward called twice through the same graph")
        grads = explicit.rasterize_bwd(ectx, None if v_image is None else v_image.contiguous(),
                                       None if v_variance is None else v_variance.contiguous())
        has_beta, has_unc = ctx.has
        return (None, grads.means, grads.quats, grads.scales, grads.densities,
                grads.beta if has_beta else None, grads.uncertainty if has_unc else None)


class _Voxelize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, options, means, quats, scales, densities, beta, uncertainty):
        output, ectx = explicit.voxelize_fwd(means.detach(), quats.detach(), scales.detach(), densities.detach(),
                                             beta=None if beta is None else beta.detach(),
                                             uncertainty=None if uncertainty is None else uncertainty.detach(),
                                             **options)
        ctx.explicit = ectx
        ctx.has = (beta is not None, uncertainty is not None)
        return output

    @staticmethod
    def backward(ctx, *grad_outputs):
        v_volume = grad_outputs[0]
        v_variance = grad_outputs[1] if len(grad_outputs) > 1 else None
        ectx, ctx.explicit = ctx.explicit, None
        if ectx is None:
            raise RuntimeError("voxelize: backward called twice through the same graph")
        grads = explicit.voxelize_bwd(ectx, None if v_volume is None else v_volume.contiguous(),
                                      None if v_variance is None else v_variance.contiguous())
        has_beta, has_unc = ctx.has
        return (None, grads.means, grads.quats, grads.scales, grads.densities,
                grads.beta if has_beta else None, grads.uncertainty if has_unc else None)


def rasterize(
    means: Tensor,
    quats: Tensor,
    scales: Tensor,
    densities: Tensor,
    viewmats: Tensor,
    Ks: Tensor,
    image_size: Sequence[int],
    *,
    beta: Optional[Tensor] = None,
    uncertainty: Optional[Tensor] = None,
    eps: float = 1.0e-8,
    radius_clip: Optional[float] = None,
    detach_variance: bool = False,
    backward_mode: str = "primitive",
    recompute: Optional[bool] = None,
    recompute_budget: Optional[int] = None,
):
    """Differentiable ``rasterize_fwd``: images [C, H, W] or (mean, variance)."""
    rc = _should_recompute(recompute, recompute_budget, means.shape[0], viewmats.shape[0], means.device)
    options = dict(viewmats=viewmats, Ks=Ks, image_size=image_size, eps=eps, radius_clip=radius_clip,
                   detach_variance=detach_variance, backward_mode=backward_mode, recompute=rc)
    return _Rasterize.apply(options, means, quats, scales, densities, beta, uncertainty)


def voxelize(
    means: Tensor,
    quats: Tensor,
    scales: Tensor,
    densities: Tensor,
    volume_shape: Sequence[int],
    *,
    volume_extent,
    volume_center=(0.0, 0.0, 0.0),
    beta: Optional[Tensor] = None,
    uncertainty: Optional[Tensor] = None,
    eps: float = 1.0e-8,
    radius_clip: Optional[float] = None,
    detach_variance: bool = False,
    backward_mode: str = "primitive",
    recompute: Optional[bool] = None,
    recompute_budget: Optional[int] = None,
):
    """Differentiable ``voxelize_fwd``: volume [D, H, W] or (mean, variance)."""
    rc = _should_recompute(recompute, recompute_budget, means.shape[0], 1, means.device)
    options = dict(volume_shape=volume_shape, volume_extent=volume_extent, volume_center=volume_center, eps=eps,
                   radius_clip=radius_clip, detach_variance=detach_variance, backward_mode=backward_mode,
                   recompute=rc)
    return _Voxelize.apply(options, means, quats, scales, densities, beta, uncertainty)


class _MSE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, target):
        from .hip import _backend as be

        if image.dtype != torch.float32 or target.dtype != torch.float32 or not image.is_cuda:
            raise TypeError("mse_loss takes float32 CUDA tensors")
        if image.shape != target.shape or image.device != target.device:
            raise ValueError(f"mse_loss: image {tuple(image.shape)} and target {tuple(target.shape)} differ")
        image, target = image.contiguous(), target.contiguous()
        n = image.numel()
        v = torch.empty_like(image)
        partials = torch.empty((max(1, int(be.lib().r2s_mse_partials(n))),), dtype=torch.float32,
                               device=image.device)
        loss = torch.empty((), dtype=torch.float32, device=image.device)
        with torch.cuda.device(image.device):
            be.call("r2s_mse_adjoint", be.ptr(image), be.ptr(target), n, be.ptr(v), be.ptr(partials), be.ptr(loss),
                    be.stream_of(image.device))
        ctx.save_for_backward(v)
        ctx.target_grad = target.requires_grad
        return loss

    @staticmethod
    def backward(ctx, grad):
        (v,) = ctx.saved_tensors
        g = v * grad
        return g, (-g if ctx.target_grad else None)


def mse_loss(image: Tensor, target: Tensor) -> Tensor:
    """mean((image - target)^2) of float32 CUDA tensors of one shape, with
    d/d image = 2 (image - target) / n from the same kernel."""
    return _MSE.apply(image, target)


__all__ = ["mse_loss", "rasterize", "voxelize"]

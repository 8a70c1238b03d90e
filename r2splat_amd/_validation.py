"""Input validation for the explicit operators (semantics of r2splat/_validation.py:14-96).

Errors raised here happen before any kernel launch: ValueError for bad values
and shapes, TypeError for bad types. Value checks read tensor contents, which
synchronises with the GPU; during hipGraph capture (where a host read is not
allowed) only the shape, dtype and device checks run, so a captured call must
have been validated by an eager call with the same inputs first.
"""

from __future__ import annotations

import math
from numbers import Integral, Real
from typing import Optional, Sequence, Union

import torch

Tensor = torch.Tensor
INT32_MAX = 2**31 - 1


def _value_checks() -> bool:
    """False while the current stream is being captured into a graph."""
    return not (torch.cuda.is_available() and torch.cuda.is_current_stream_capturing())


def _shape(name: str, value: Sequence[int], rank: int) -> tuple:
    if not isinstance(value, (tuple, list, torch.Size)) or len(value) != rank:
        raise ValueError(f"{name} must contain {rank} positive integers")
    for v in value:
        if isinstance(v, bool) or not isinstance(v, Integral) or v <= 0:
            raise ValueError(f"{name} must contain {rank} positive integers")
    out = tuple(int(v) for v in value)
    if math.prod(out) > INT32_MAX:
        raise ValueError(f"{name} exceeds the kernels' signed 32-bit indexing limit")
    return out


def _scalar(name: str, value, *, positive: bool = False) -> float:
    if isinstance(value, bool) or not isinstance(value, Real):
        raise TypeError(f"{name} must be a finite number")
    value = float(value)
    bad = value <= 0 if positive else value < 0
    if not math.isfinite(value) or bad:
        raise ValueError(f"{name} must be finite and {'positive' if positive else 'nonnegative'}")
    return value


def _tensor(name: str, value, shape: tuple, reference: Optional[Tensor] = None) -> Tensor:
    if not isinstance(value, Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if value.ndim != len(shape) or any(s is not None and s != d for s, d in zip(shape, value.shape)):
        raise ValueError(f"{name} must have shape {shape}; got {tuple(value.shape)}")
    if value.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"{name} must have dtype torch.float32 or torch.float64")
    if not value.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if reference is not None and (value.device != reference.device or value.dtype != reference.dtype):
        raise ValueError(f"{name} must have the same device and dtype as means")
    if _value_checks() and not bool(torch.isfinite(value).all()):
        raise ValueError(f"{name} must contain only finite values")
    return value.contiguous()


def _constant(name: str, value: Tensor) -> Tensor:
    if value.requires_grad:
        raise ValueError(f"gradients with respect to {name} are not supported")
    return value


def _vector(name: str, value: Union[Tensor, Sequence[float]], reference: Tensor) -> Tensor:
    if not isinstance(value, Tensor):
        value = torch.as_tensor(value, device=reference.device, dtype=reference.dtype)
    return _constant(name, _tensor(name, value, (3,), reference))


def _primitives(means, quats, scales, densities, beta, uncertainty):
    means = _tensor("means", means, (None, 3))
    n = means.shape[0]
    if n > INT32_MAX:
        raise ValueError("primitive count exceeds the kernels' signed 32-bit indexing limit")
    quats = _tensor("quats", quats, (n, 4), means)
    scales = _tensor("scales", scales, (n, 3), means)
    densities = _tensor("densities", densities, (n,), means)
    checks = _value_checks()
    if checks and bool((quats == 0).all(dim=-1).any()):
        raise ValueError("quats must have nonzero norm")
    if checks and bool((scales <= 0).any()):
        raise ValueError("scales must be strictly positive physical lengths")
    if checks and bool((densities < 0).any()):
        raise ValueError("densities must be nonnegative")
    if beta is not None:
        beta = _tensor("beta", beta, (n,), means)
        if checks and bool((beta <= 0).any()):
            raise ValueError("beta must contain positive 3D exponents")
    if uncertainty is not None:
        uncertainty = _tensor("uncertainty", uncertainty, (n,), means)
        if checks and bool((uncertainty < 0).any()):
            raise ValueError("uncertainty must be nonnegative relative variance")
    return means, quats, scales, densities, beta, uncertainty


def _options(beta, eps, radius_clip, detach_variance):
    if not isinstance(detach_variance, bool):
        raise TypeError("detach_variance must be a bool")
    eps = _scalar("eps", eps)
    if radius_clip is None:
        radius_clip = 3.0 if beta is None else 1.0
    return eps, _scalar("radius_clip", radius_clip, positive=True)

"""Argument checking for the explicit operators.

Behaviour (error classes, message wording, defaults) follows the reference's
r2splat/_validation.py:14-96 and the preambles of r2splat/explicit.py:118-160;
the structure is this package's own. Checks run in two phases:

* structural checks (type, rank, shape, dtype, device) read only tensor
  metadata and raise as each argument is read;
* value checks (finiteness and each field's domain, such as positive scales)
  are collected as device-side predicates in a ``ValueChecks`` batch and
  evaluated with a single device-to-host copy per call, then reported in a
  fixed precedence order. While the current stream is being captured into a
  graph (where a host read is not allowed) the batch is skipped, so a
  captured call must have been validated by an eager call with the same
  inputs first.

TypeError reports a wrong kind of argument, ValueError a wrong value or shape.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from numbers import Integral, Real
from typing import Callable, List, Optional, Sequence, Tuple

import torch

Tensor = torch.Tensor
INT32_MAX = 2**31 - 1
FLOATS = (torch.float32, torch.float64)


def capturing() -> bool:
    """True while the current stream is being captured into a graph."""
    return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()


# ---------------------------------------------------------------------------
# deferred value checks
# ---------------------------------------------------------------------------
class ValueChecks:
    """Device-side predicates ("some element is bad") evaluated together.

    ``run()`` stacks every predicate into one boolean vector, copies it to the
    host once and raises ValueError with the message of the first predicate
    that holds, in the order they were added.
    """

    def __init__(self):
        self._pending: List[Tuple[str, Callable[[], Tensor]]] = []

    def add(self, message: str, any_bad: Callable[[], Tensor]) -> None:
        self._pending.append((message, any_bad))

    def finite(self, name: str, value: Tensor) -> None:
        self.add(f"{name} must contain only finite values", lambda: ~torch.isfinite(value).all())

    def run(self) -> None:
        pending, self._pending = self._pending, []
        if not pending or capturing():
            return
        flags = torch.stack([fn().reshape(()) for _, fn in pending]).tolist()
        for (message, _), bad in zip(pending, flags):
            if bad:
                raise ValueError(message)


# ---------------------------------------------------------------------------
# structural checks
# ---------------------------------------------------------------------------
def positive_ints(name: str, value: Sequence[int], rank: int) -> tuple:
    """An output shape: `rank` positive Python integers whose product fits int32."""
    ok = isinstance(value, (tuple, list, torch.Size)) and len(value) == rank and all(
        isinstance(v, Integral) and not isinstance(v, bool) and v > 0 for v in value)
    if not ok:
        raise ValueError(f"{name} must contain {rank} positive integers")
    dims = tuple(int(v) for v in value)
    if math.prod(dims) > INT32_MAX:
        raise ValueError(f"{name} exceeds the kernels' signed 32-bit indexing limit")
    return dims


def number(name: str, value, *, positive: bool = False) -> float:
    """A finite real option, nonnegative (or positive)."""
    if isinstance(value, bool) or not isinstance(value, Real):
        raise TypeError(f"{name} must be a finite number")
    x = float(value)
    lower_ok = x > 0 if positive else x >= 0
    if not (math.isfinite(x) and lower_ok):
        raise ValueError(f"{name} must be finite and {'positive' if positive else 'nonnegative'}")
    return x


def tensor(name: str, value, shape: tuple, like: Optional[Tensor] = None) -> Tensor:
    """A floating GPU tensor of `shape` (None = any extent) with `like`'s device
    and dtype, made contiguous. Finiteness is a value check (ValueChecks.finite)."""
    if not isinstance(value, Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if value.ndim != len(shape) or any(want is not None and want != got for want, got in zip(shape, value.shape)):
        raise ValueError(f"{name} must have shape {shape}; got {tuple(value.shape)}")
    if value.dtype not in FLOATS:
        raise TypeError(f"{name} must have dtype torch.float32 or torch.float64")
    if not value.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if like is not None and (value.device, value.dtype) != (like.device, like.dtype):
        raise ValueError(f"{name} must have the same device and dtype as means")
    return value.contiguous()


def constant(name: str, value) -> None:
    """Geometry is constant: no gradient flows to it."""
    if isinstance(value, Tensor) and value.requires_grad:
        raise ValueError(f"gradients with respect to {name} are not supported")


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class _Field:
    name: str
    trailing: tuple  # dims after the primitive count
    required: bool = True
    violation: Optional[Callable[[Tensor], Tensor]] = None  # elementwise "bad value"
    message: str = ""


_FIELDS = (
    _Field("means", (3,)),
    _Field("quats", (4,), True, lambda t: (t == 0).all(dim=-1), "quats must have nonzero norm"),
    _Field("scales", (3,), True, lambda t: t <= 0, "scales must be strictly positive physical lengths"),
    _Field("densities", (), True, lambda t: t < 0, "densities must be nonnegative"),
    _Field("beta", (), False, lambda t: t <= 0, "beta must contain positive 3D exponents"),
    _Field("uncertainty", (), False, lambda t: t < 0, "uncertainty must be nonnegative relative variance"),
)


def _domain(checks: ValueChecks, field: _Field, value: Tensor) -> None:
    if field.violation is not None:
        checks.add(field.message, lambda: field.violation(value).any())


def primitives(values: Sequence, checks: ValueChecks) -> tuple:
    """(means, quats, scales, densities, beta, uncertainty), the last two
    optional (None): structural checks now, value checks queued on `checks`
    (reported as the reference does: the required fields' finiteness, their
    domains, then each optional field's finiteness and domain)."""
    got = {}
    like = None
    for field, value in zip(_FIELDS, values):
        if value is None and not field.required:
            got[field.name] = None
            continue
        lead = (None,) if like is None else (like.shape[0],)
        got[field.name] = tensor(field.name, value, lead + field.trailing, like)
        if like is None:
            like = got[field.name]
            if like.shape[0] > INT32_MAX:
                raise ValueError("primitive count exceeds the kernels' signed 32-bit indexing limit")
    required = [f for f in _FIELDS if f.required]
    for f in required:
        checks.finite(f.name, got[f.name])
    for f in required:
        _domain(checks, f, got[f.name])
    for f in _FIELDS:
        if not f.required and got[f.name] is not None:
            checks.finite(f.name, got[f.name])
            _domain(checks, f, got[f.name])
    return tuple(got[f.name] for f in _FIELDS)


def options(beta, eps, radius_clip, detach_variance) -> tuple:
    """(eps, radius_clip) with the reference defaults: radius_clip 3 for
    Gaussian and 1 for Beta primitives."""
    if not isinstance(detach_variance, bool):
        raise TypeError("detach_variance must be a bool")
    eps = number("eps", eps)
    clip = (3.0 if beta is None else 1.0) if radius_clip is None else radius_clip
    return eps, number("radius_clip", clip, positive=True)

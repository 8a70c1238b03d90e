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

import ctypes
import functools
import math
from dataclasses import dataclass
from numbers import Integral, Real
from typing import Callable, List, Optional, Sequence, Tuple

import torch

from .hip import _backend as be

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
    """Predicates ("some element is bad") evaluated together.

    Elementwise checks of a whole buffer (finiteness, sign, all-zero rows)
    run as one native launch per 16 (r2s_value_checks); other predicates are
    torch reductions; results already known on the host keep their place in
    the order. ``run()`` copies every flag to the host once and raises
    ValueError with the message of the first predicate that holds, in the
    order they were added.
    """

    def __init__(self):
        # (message, check): check is a callable (torch predicate), a bool
        # (known on the host) or (tensor, kind, row) (native elementwise)
        self._pending: List[Tuple[str, object]] = []

    def add(self, message: str, any_bad: Callable[[], Tensor]) -> None:
        self._pending.append((message, any_bad))

    def known(self, message: str, bad: bool) -> None:
        self._pending.append((message, bool(bad)))

    def element(self, message: str, value: Tensor, kind: int, row: int = 1) -> None:
        """Some element of `value` (or some row of `row` elements, kind
        CHECK_ZERO_ROW) is bad; `kind` one of the CHECK_* codes."""
        self._pending.append((message, (value.contiguous(), kind, row)))

    def finite(self, name: str, value: Tensor) -> None:
        self.element(f"{name} must contain only finite values", value, be.CHECK_NONFINITE)

    def run(self) -> None:
        pending, self._pending = self._pending, []
        if not pending or capturing():
            return
        native = [i for i, (_, c) in enumerate(pending) if isinstance(c, tuple)]
        device = pending[native[0]][1][0].device if native else None
        for i in list(native):
            value, kind, row = pending[i][1]
            if value.device != device or value.dtype not in FLOATS:
                native.remove(i)
                pending[i] = (pending[i][0], functools.partial(_element_predicate, value, kind, row))
        reduced = [i for i, (_, c) in enumerate(pending) if callable(c)]
        parts = []
        if native:
            flags = torch.empty((len(native),), dtype=torch.int32, device=device)
            with torch.cuda.device(device):
                for lo in range(0, len(native), be.MAX_VALUE_CHECKS):
                    chunk = native[lo: lo + be.MAX_VALUE_CHECKS]
                    table = (be.ValueCheck * len(chunk))()
                    for entry, i in zip(table, chunk):
                        value, kind, row = pending[i][1]
                        entry.data, entry.n, entry.dtype = value.data_ptr(), value.numel(), be.dtype_code(value.dtype)
                        entry.kind, entry.row = kind, row
                    be.call("r2s_value_checks", ctypes.addressof(table), len(chunk), flags[lo:].data_ptr(),
                            be.stream_of(device))
            parts.append(flags)
        if reduced:
            parts.append(torch.stack([pending[i][1]().reshape(()) for i in reduced]).to(torch.int32))
        values = (torch.cat(parts) if len(parts) > 1 else parts[0]).tolist() if parts else []
        bad = dict(zip(native + reduced, values))
        for i, (message, c) in enumerate(pending):
            if (c if isinstance(c, bool) else bad[i]):
                raise ValueError(message)


def _element_predicate(value: Tensor, kind: int, row: int) -> Tensor:
    """The torch form of a native elementwise check (buffers the native launch
    cannot take: another device or dtype)."""
    if kind == be.CHECK_NONFINITE:
        return ~torch.isfinite(value).all()
    if kind == be.CHECK_NONPOSITIVE:
        return (value <= 0).any()
    if kind == be.CHECK_NEGATIVE:
        return (value < 0).any()
    m = value.reshape(-1, row)
    if kind == be.CHECK_K_FOCAL:
        return (m[:, (0, 4)] <= 0).any()
    if kind == be.CHECK_K_PINHOLE:
        return (m[:, (1, 3, 6, 7)] != 0).any() | (m[:, 8] != 1).any()
    if kind == be.CHECK_VIEW_ROW:
        return (m[:, 12:15] != 0).any() | (m[:, 15] != 1).any()
    return (m == 0).all(dim=-1).any()


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
    violation: Optional[int] = None  # elementwise "bad value": a CHECK_* code of r2s_value_checks
    message: str = ""


_FIELDS = (
    _Field("means", (3,)),
    _Field("quats", (4,), True, be.CHECK_ZERO_ROW, "quats must have nonzero norm"),
    _Field("scales", (3,), True, be.CHECK_NONPOSITIVE, "scales must be strictly positive physical lengths"),
    _Field("densities", (), True, be.CHECK_NEGATIVE, "densities must be nonnegative"),
    _Field("beta", (), False, be.CHECK_NONPOSITIVE, "beta must contain positive 3D exponents"),
    _Field("uncertainty", (), False, be.CHECK_NEGATIVE, "uncertainty must be nonnegative relative variance"),
)


def _domain(checks: ValueChecks, field: _Field, value: Tensor) -> None:
    if field.violation is not None:
        checks.element(field.message, value, field.violation, field.trailing[0] if field.trailing else 1)


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

"""Line-integral normalisation of compact Beta primitives.

A 3D Beta primitive density (1 - d^2)_+^gamma integrates along a ray through
its centre to B(1/2, gamma + 1) times the ray half-length; relative to the
Gaussian path factor sqrt(2 pi) * half-length this is the ratio

    r(gamma) = B(1/2, gamma + 1) / sqrt(2 pi) = Gamma(gamma + 1) / (sqrt(2) Gamma(gamma + 3/2)).

Semantics follow r2splat/_beta.py:10-49. Small exponents use lgamma/digamma in
float64; from gamma >= 20 the Stirling-series difference
log Gamma(z + 1/2) - log Gamma(z) = 1/2 log z + sum_k (2^-k - 2) B_(k+1) / (k (k+1) z^k)
(z = gamma + 1, odd k, Bernoulli numbers B_n) is used, which avoids the
cancellation of two large lgamma values.
"""

from __future__ import annotations

import math

import torch

_SWITCH = 20.0
# (power of 1/z, coefficient) of log Gamma(z) - log Gamma(z + 1/2) + 1/2 log z.
_SERIES = ((1, 1.0 / 8.0), (3, -1.0 / 192.0), (5, 1.0 / 640.0), (7, -17.0 / 14336.0), (9, 31.0 / 18432.0))


def _log_ratio(value: torch.Tensor) -> torch.Tensor:
    small = value.clamp_max(_SWITCH)
    direct = torch.lgamma(small + 1.0) - torch.lgamma(small + 1.5) - 0.5 * math.log(2.0)
    z = value.clamp_min(_SWITCH) + 1.0
    w = z.reciprocal()
    series = torch.zeros_like(z)
    for power, coefficient in _SERIES:
        series = series + coefficient * w.pow(power)
    asymptotic = series - 0.5 * (math.log(2.0) + z.log())
    return torch.where(value < _SWITCH, direct, asymptotic)


def beta_path_ratio(gamma: torch.Tensor) -> torch.Tensor:
    """B(1/2, gamma + 1) / sqrt(2 pi), evaluated in float64 and cast back."""
    return _log_ratio(gamma.to(torch.float64)).exp().to(gamma.dtype)


def beta_path_log_grad(gamma: torch.Tensor) -> torch.Tensor:
    """d log B(1/2, gamma + 1) / d gamma = psi(gamma + 1) - psi(gamma + 3/2)."""
    value = gamma.to(torch.float64)
    small = value.clamp_max(_SWITCH)
    direct = torch.special.digamma(small + 1.0) - torch.special.digamma(small + 1.5)
    z = value.clamp_min(_SWITCH) + 1.0
    w = z.reciprocal()
    derivative = -0.5 * w
    for power, coefficient in _SERIES:
        derivative = derivative - power * coefficient * w.pow(power + 1)
    return torch.where(value < _SWITCH, direct, derivative).to(gamma.dtype)


__all__ = ["beta_path_ratio", "beta_path_log_grad"]

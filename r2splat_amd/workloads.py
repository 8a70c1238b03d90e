"""Synthetic CT scenes of the benchmark configurations (no datasets needed).

Primitives: means ~ U[-0.5, 0.5]^3, quats ~ N(0, 1)^4 (unnormalised), scales
log-uniform per axis, densities ~ U[0, 1]. Cameras: C views evenly spaced over
360 degrees on a circle of radius 2.5 in the x-z plane, looking at the origin,
world +y up, OpenCV frame (x right, y down, z forward); fx = fy = 1.2 W,
cx = W / 2, cy = H / 2.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch


@dataclass
class Scene:
    means: torch.Tensor
    quats: torch.Tensor
    scales: torch.Tensor
    densities: torch.Tensor
    viewmats: torch.Tensor
    Ks: torch.Tensor
    height: int
    width: int
    beta: torch.Tensor | None = None
    uncertainty: torch.Tensor | None = None

    @property
    def image_size(self):
        return (self.height, self.width)

    def params(self):
        return self.means, self.quats, self.scales, self.densities


def orbit_cameras(n_views: int, width: int, height: int, radius: float = 2.5, dtype=torch.float32,
                  device="cpu"):
    views, Ks = [], []
    up = torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64)
    for k in range(n_views):
        theta = 2.0 * math.pi * k / n_views
        centre = torch.tensor([radius * math.sin(theta), 0.0, -radius * math.cos(theta)], dtype=torch.float64)
        z = -centre / centre.norm()
        x = torch.linalg.cross(-up, z)
        x = x / x.norm()
        y = torch.linalg.cross(z, x)
        R = torch.stack([x, y, z])
        V = torch.eye(4, dtype=torch.float64)
        V[:3, :3] = R
        V[:3, 3] = -R @ centre
        views.append(V)
        f = 1.2 * width
        Ks.append(torch.tensor([[f, 0.0, width / 2.0], [0.0, 1.2 * width, height / 2.0], [0.0, 0.0, 1.0]],
                               dtype=torch.float64))
    return torch.stack(views).to(device, dtype), torch.stack(Ks).to(device, dtype)


def primitives(n: int, scale_range=(0.005, 0.05), seed: int = 0, dtype=torch.float32, device="cpu",
               beta_range=None, uncertainty_max=None):
    g = torch.Generator().manual_seed(seed)
    means = torch.rand(n, 3, generator=g, dtype=torch.float64) - 0.5
    quats = torch.randn(n, 4, generator=g, dtype=torch.float64)
    lo, hi = math.log(scale_range[0]), math.log(scale_range[1])
    scales = torch.exp(lo + (hi - lo) * torch.rand(n, 3, generator=g, dtype=torch.float64))
    densities = torch.rand(n, generator=g, dtype=torch.float64)
    out = [means, quats, scales, densities]
    beta = unc = None
    if beta_range is not None:
        bl, bh = math.log(beta_range[0]), math.log(beta_range[1])
        beta = torch.exp(bl + (bh - bl) * torch.rand(n, generator=g, dtype=torch.float64))
    if uncertainty_max is not None:
        unc = uncertainty_max * torch.rand(n, generator=g, dtype=torch.float64)
    conv = [t.to(device, dtype) for t in out]
    return (*conv, None if beta is None else beta.to(device, dtype), None if unc is None else unc.to(device, dtype))


def scene(n: int, n_views: int, height: int, width: int, *, scale_range=(0.005, 0.05), seed: int = 0,
          dtype=torch.float32, device="cpu", beta_range=None, uncertainty_max=None) -> Scene:
    means, quats, scales, dens, beta, unc = primitives(n, scale_range, seed, dtype, device, beta_range,
                                                       uncertainty_max)
    views, Ks = orbit_cameras(n_views, width, height, dtype=dtype, device=device)
    return Scene(means, quats, scales, dens, views, Ks, height, width, beta, unc)


def config0(device="cpu", dtype=torch.float32, seed=0) -> Scene:
    """N = 1,000, C = 4 views of 64 x 64, scales log-uniform in [0.02, 0.10]."""
    return scene(1000, 4, 64, 64, scale_range=(0.02, 0.10), seed=seed, dtype=dtype, device=device)


def config1(device="cpu", dtype=torch.float32, seed=0) -> Scene:
    """N = 100,000, C = 64 views of 256 x 256, scales log-uniform in [0.005, 0.05]."""
    return scene(100_000, 64, 256, 256, scale_range=(0.005, 0.05), seed=seed, dtype=dtype, device=device)

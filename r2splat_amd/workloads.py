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


def config2(device="cpu", dtype=torch.float32, seed=0) -> Scene:
    """N = 500,000 Beta primitives (beta log-uniform in [0.5, 4]) with
    uncertainty uniform in [0, 0.1], scales log-uniform in [0.003, 0.03],
    C = 256 views of 512 x 512."""
    return scene(500_000, 256, 512, 512, scale_range=(0.003, 0.03), seed=seed, dtype=dtype, device=device,
                 beta_range=(0.5, 4.0), uncertainty_max=0.1)


def config3_specs(n_jobs: int = 32, seed: int = 0):
    """Batched uneven reconstructions: per job N log-uniform in [10k, 400k],
    C uniform in {16, 32, 64, 128} views and image side uniform in
    {128, 256, 512}; returns [(N, C, side, job_seed)]."""
    g = torch.Generator().manual_seed(seed)
    specs = []
    for j in range(n_jobs):
        u = float(torch.rand((), generator=g, dtype=torch.float64))
        n = int(round(math.exp(math.log(10_000) + u * (math.log(400_000) - math.log(10_000)))))
        c = (16, 32, 64, 128)[int(torch.randint(0, 4, (), generator=g))]
        side = (128, 256, 512)[int(torch.randint(0, 3, (), generator=g))]
        specs.append((n, c, side, 1000 + 2 * j))
    return specs


def fit_job(n: int, views: int, side: int, job_seed: int, device="cuda", scale_range=(0.005, 0.05),
            beta_range=None):
    """One reconstruction job as in config 1: parameters from `job_seed`, target
    images rendered from a second primitive set (`job_seed + 1`); Beta
    primitives (betas uniform in `beta_range`) when it is given."""
    from .explicit import rasterize_fwd

    sc = scene(n, views, side, side, scale_range=scale_range, seed=job_seed, device=device, beta_range=beta_range)
    tg = scene(n, views, side, side, scale_range=scale_range, seed=job_seed + 1, device=device,
               beta_range=beta_range)
    with torch.no_grad():
        target, _ = rasterize_fwd(*tg.params(), tg.viewmats, tg.Ks, tg.image_size, beta=tg.beta)
    return dict(means=sc.means, quats=sc.quats, scales=sc.scales, densities=sc.densities, beta=sc.beta,
                viewmats=sc.viewmats, Ks=sc.Ks, target=target)

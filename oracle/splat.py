"""Float64 CPU oracle of the CT splatting operators — TEST INFRASTRUCTURE ONLY.

Differentiable PyTorch (CPU, float64) restatement of the reference mathematics,
written independently of the HIP kernels (dense matrix formulation, library
inverses), plus exact integer tile binning in numpy. Every function cites the
reference file:line it restates (paths inside the reference repository).

Pinned by tests/test_oracle.py against: the reference's golden tables
(Beta path factor, 3D tile-id known answers), closed forms (centred Gaussian
and Beta line integrals, 1x1 images), independent formulations (quaternion
cross-product covariance), invariants (quaternion scale invariance, symmetry)
and brute force on tiny grids.
"""

from __future__ import annotations

import math
from typing import Optional

import numpy as np
import torch

Tensor = torch.Tensor
TWO_PI = 2.0 * math.pi


# ---------------------------------------------------------------------------
# packing helpers
# ---------------------------------------------------------------------------
_PACK = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def pack_sym(m: Tensor) -> Tensor:
    """[...,3,3] -> [...,6] as (00,01,02,11,12,22)."""
    return torch.stack([m[..., i, j] for i, j in _PACK], dim=-1)


def unpack_sym(p: Tensor) -> Tensor:
    """[...,6] -> symmetric [...,3,3]."""
    rows = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    return torch.stack([torch.stack([p[..., k] for k in r], dim=-1) for r in rows], dim=-2)


# ---------------------------------------------------------------------------
# parameter transform (r2splat/cuda/csrc/ParameterTransformFwd.cu:10-40,
# r2splat/cuda/include/ParameterTransform.cuh:111-140)
# ---------------------------------------------------------------------------
def rotation(quats: Tensor) -> Tensor:
    """Rotation matrices [N,3,3] of unnormalised (w,x,y,z) quaternions."""
    q = quats / torch.linalg.vector_norm(quats, dim=-1, keepdim=True)
    w, x, y, z = q.unbind(-1)
    return torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def covariance(quats: Tensor, scales: Tensor) -> Tensor:
    """R diag(s^2) R^T, [N,3,3]."""
    R = rotation(quats)
    return R @ torch.diag_embed(scales.square()) @ R.transpose(-1, -2)


def covariance_cross_product(quats: Tensor, scales: Tensor) -> Tensor:
    """Independent form: rotate the basis with v' = v + 2 w (u x v) + 2 u x (u x v)."""
    q = quats / torch.linalg.vector_norm(quats, dim=-1, keepdim=True)
    basis = torch.eye(3, dtype=q.dtype).expand(q.shape[0], 3, 3)
    u = q[:, None, 1:].expand_as(basis)
    c1 = torch.linalg.cross(u, basis, dim=-1)
    rotated = basis + 2 * (q[:, :1, None] * c1 + torch.linalg.cross(u, c1, dim=-1))
    R = rotated.transpose(-1, -2)
    return R @ torch.diag_embed(scales.square()) @ R.transpose(-1, -2)


# ---------------------------------------------------------------------------
# world -> image projection (WorldToImageProjectionFwd.cu:73-283)
# ---------------------------------------------------------------------------
def project(means: Tensor, cov3: Tensor, viewmats: Tensor, Ks: Tensor, width: int, height: int):
    """Returns (means2d [C,N,2], cov2d [C,N,3,3], visible [C,N]).

    cov2d = J Rc Sigma Rc^T J^T with the local perspective Jacobian whose third
    row is the unit ray direction; the ray point is clamped to 15% outside the
    image in normalised coordinates. Culled (z <= 0.01) entries are zero.
    """
    Rc, t = viewmats[:, :3, :3], viewmats[:, :3, 3]
    pc = torch.einsum("cij,nj->cni", Rc, means) + t[:, None]
    x, y, z = pc.unbind(-1)
    visible = z > 0.01
    z = torch.where(visible, z, torch.ones_like(z))
    fx, fy = Ks[:, 0, 0, None], Ks[:, 1, 1, None]
    cx, cy = Ks[:, 0, 2, None], Ks[:, 1, 2, None]
    u = torch.minimum(torch.maximum(x / z, -(cx + 0.15 * width) / fx), (1.15 * width - cx) / fx)
    v = torch.minimum(torch.maximum(y / z, -(cy + 0.15 * height) / fy), (1.15 * height - cy) / fy)
    tx, ty = z * u, z * v
    zero = torch.zeros_like(z)
    ray = torch.stack([tx, ty, z], -1)
    J = torch.stack([
        torch.stack([fx / z, zero, -fx * tx / z**2], -1),
        torch.stack([zero, fy / z, -fy * ty / z**2], -1),
        ray / torch.linalg.vector_norm(ray, dim=-1, keepdim=True),
    ], -2)
    Sc = Rc[:, None] @ cov3[None] @ Rc[:, None].transpose(-1, -2)
    cov2d = J @ Sc @ J.transpose(-1, -2)
    means2d = torch.stack([fx * x / z + cx, fy * y / z + cy], -1)
    vis = visible[..., None]
    return torch.where(vis, means2d, 0.0), torch.where(vis[..., None], cov2d, 0.0), visible


# ---------------------------------------------------------------------------
# conics and path factor (ConicsMuFwd.cu:35-104)
# ---------------------------------------------------------------------------
def conics_mu(cov2d: Tensor, eps: float):
    """(Sigma_2x2 + eps I)^-1 packed (00,01,11) and mu = sqrt(2 pi s), s the Schur
    complement of the eps-regularised 3x3; zeros where the 2x2 block is not PD."""
    A = cov2d[..., :2, :2] + eps * torch.eye(2, dtype=cov2d.dtype)
    det = A[..., 0, 0] * A[..., 1, 1] - A[..., 0, 1] * A[..., 1, 0]
    ok = (A[..., 0, 0] > 0) & (A[..., 1, 1] > 0) & (det > 0)
    safe = torch.where(ok[..., None, None], A, torch.eye(2, dtype=cov2d.dtype).expand_as(A))
    P = torch.linalg.inv(safe)
    cvec = cov2d[..., :2, 2]
    s = cov2d[..., 2, 2] + eps - (cvec * (P @ cvec[..., None]).squeeze(-1)).sum(-1)
    s_ok = ok & (s > 0)
    mu = torch.where(s_ok, torch.sqrt(TWO_PI * torch.where(s_ok, s, torch.ones_like(s))), 0.0)
    conics = torch.stack([P[..., 0, 0], P[..., 0, 1], P[..., 1, 1]], -1)
    return torch.where(ok[..., None], conics, 0.0), mu


def beta_ratio(beta: Tensor) -> Tensor:
    """B(1/2, beta + 1) / sqrt(2 pi) from the Gamma-function identity (moderate beta)."""
    return torch.exp(math.lgamma(0.5) + torch.lgamma(beta + 1) - torch.lgamma(beta + 1.5)) / math.sqrt(TWO_PI)


# ---------------------------------------------------------------------------
# radii and tile binning (WorldToImageProjectionFwd.cu:250-282,
# IsectTilesKernel.cu:30-273, IsectTiles3DKernel.cu:36-276) — exact, numpy
# ---------------------------------------------------------------------------
def radii_2d(cov2d_packed: np.ndarray, means2d: np.ndarray, eps: float, clip: float, width: int, height: int):
    """Integer radii [C,N,2] computed in the dtype of the inputs (float32 or float64)."""
    dt = cov2d_packed.dtype.type
    a = cov2d_packed[..., 0] + dt(np.float32(eps))
    d = cov2d_packed[..., 3] + dt(np.float32(eps))
    b = cov2d_packed[..., 1]
    with np.errstate(invalid="ignore", over="ignore"):
        ok = (a * d - b * b) > 0
        rx = np.ceil(dt(np.float32(clip)) * np.sqrt(np.maximum(dt(0), a)))
        ry = np.ceil(dt(np.float32(clip)) * np.sqrt(np.maximum(dt(0), d)))
    ok &= np.isfinite(rx) & np.isfinite(ry) & (rx > 0) & (ry > 0)
    mx, my = means2d[..., 0], means2d[..., 1]
    ok &= ~((mx + rx <= 0) | (mx - rx >= width) | (my + ry <= 0) | (my - ry >= height))
    out = np.zeros(cov2d_packed.shape[:-1] + (2,), dtype=np.int32)
    out[..., 0] = np.where(ok, rx, 0).astype(np.int32)
    out[..., 1] = np.where(ok, ry, 0).astype(np.int32)
    return out


def _span(mean: np.ndarray, radius: np.ndarray, tile: int, n_tiles: int):
    dt = mean.dtype.type
    t = dt(tile)
    with np.errstate(invalid="ignore", over="ignore"):
        c = mean / t
        r = radius.astype(mean.dtype) / t
        lo_f, hi_f = np.floor(c - r), np.ceil(c + r)

    def clamp(v):
        out = np.where(np.isfinite(v) & (v > 0), v, 0)
        out = np.where(out >= n_tiles, n_tiles, out)
        return out.astype(np.int64)

    return clamp(lo_f), clamp(hi_f)


def tile_rects_2d(means2d: np.ndarray, radii: np.ndarray, tile: int, ntx: int, nty: int):
    """(x0, x1, y0, y1) half-open tile ranges per (c, n); empty when radii <= 0."""
    rx, ry = radii[..., 0], radii[..., 1]
    x0, x1 = _span(means2d[..., 0], rx, tile, ntx)
    y0, y1 = _span(means2d[..., 1], ry, tile, nty)
    valid = (rx > 0) & (ry > 0) & np.isfinite(means2d[..., 0]) & np.isfinite(means2d[..., 1])
    valid &= (x1 > x0) & (y1 > y0)
    for a in (x0, x1, y0, y1):
        a[~valid] = 0
    return x0, x1, y0, y1


def _emit(counts: np.ndarray, lo: list, widths: list, strides: list, base: np.ndarray):
    """Row-major enumeration of box cells per owner; returns (keys, owners)."""
    owners = np.repeat(np.arange(counts.size, dtype=np.int64), counts)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    local = np.arange(owners.size, dtype=np.int64) - np.repeat(first, counts)
    keys = np.repeat(base, counts).astype(np.int64)
    # innermost dimension is last in `widths`
    rem = local
    for k in reversed(range(len(widths))):
        w = np.repeat(widths[k], counts)
        coord = rem % np.maximum(w, 1) + np.repeat(lo[k], counts)
        rem = rem // np.maximum(w, 1)
        keys += coord * strides[k]
    return keys, owners


def isect_tiles_2d(means2d: np.ndarray, radii: np.ndarray, tile: int, ntx: int, nty: int):
    """(tiles_per_gauss [C,N], isect_ids [I] int64, flatten_ids [I] int32), stable by key."""
    C, N = means2d.shape[:2]
    x0, x1, y0, y1 = tile_rects_2d(means2d, radii, tile, ntx, nty)
    counts = ((x1 - x0) * (y1 - y0)).reshape(-1)
    cam = np.repeat(np.arange(C, dtype=np.int64), N)
    keys, owners = _emit(counts, [y0.reshape(-1), x0.reshape(-1)], [(y1 - y0).reshape(-1), (x1 - x0).reshape(-1)],
                         [ntx, 1], cam * (ntx * nty))
    order = np.argsort(keys, kind="stable")
    return counts.reshape(C, N).astype(np.int32), keys[order], owners[order].astype(np.int32)


def isect_tiles_3d(means3d: np.ndarray, radii: np.ndarray, tile: int, ntx: int, nty: int, ntz: int):
    N = means3d.shape[0]
    lo, hi = [], []
    for k, nt in enumerate((ntx, nty, ntz)):
        a, b = _span(means3d[:, k], radii[:, k], tile, nt)
        lo.append(a)
        hi.append(b)
    valid = (radii > 0).all(axis=1)
    for k in range(3):
        valid &= hi[k] > lo[k]
    w = [np.where(valid, hi[k] - lo[k], 0) for k in range(3)]
    counts = (w[0] * w[1] * w[2]).astype(np.int64)
    keys, owners = _emit(counts, [lo[2], lo[1], lo[0]], [w[2], w[1], w[0]], [ntx * nty, ntx, 1],
                         np.zeros(N, dtype=np.int64))
    order = np.argsort(keys, kind="stable")
    return counts.astype(np.int32), keys[order], owners[order].astype(np.int32)


def offsets(isect_ids: np.ndarray, n_keys: int) -> np.ndarray:
    return np.searchsorted(isect_ids, np.arange(n_keys), side="left").astype(np.int32)


def tile_mask_2d(means2d: np.ndarray, radii: np.ndarray, tile: int, width: int, height: int) -> np.ndarray:
    """[C,N,H,W] bool: pixel receives (c, n) iff its tile lies in the (c, n) tile rect."""
    ntx, nty = -(-width // tile), -(-height // tile)
    x0, x1, y0, y1 = tile_rects_2d(means2d, radii, tile, ntx, nty)
    tx = np.arange(width) // tile
    ty = np.arange(height) // tile
    mx = (tx[None, None, :] >= x0[..., None]) & (tx[None, None, :] < x1[..., None])
    my = (ty[None, None, :] >= y0[..., None]) & (ty[None, None, :] < y1[..., None])
    return my[..., :, None] & mx[..., None, :]


def tile_mask_3d(means3d: np.ndarray, radii: np.ndarray, tile: int, shape_zyx) -> np.ndarray:
    """[N,D,H,W] bool analogue of tile_mask_2d for voxels."""
    D, H, W = shape_zyx
    masks = []
    for k, size in enumerate((W, H, D)):
        nt = -(-size // tile)
        lo, hi = _span(means3d[:, k], radii[:, k], tile, nt)
        t = np.arange(size) // tile
        masks.append((t[None] >= lo[:, None]) & (t[None] < hi[:, None]) & (radii > 0).all(axis=1)[:, None])
    mx, my, mz = masks
    return mz[:, :, None, None] & my[:, None, :, None] & mx[:, None, None, :]


# ---------------------------------------------------------------------------
# rasterization (RasterizeToPixelsFwd.cu:155-180, :664-680, mean/var :318-352)
# ---------------------------------------------------------------------------
def _camera_alpha(mx, my, conics, opac, mus, width, height, beta2d=None, mask=None, visible=None):
    """Kept alpha [N,H,W] of one camera's primitives at every pixel centre
    (RasterizeToPixelsFwd.cu:155-180 Gaussian, :664-680 Beta); mx, my
    broadcast against [N,H,W] (per-primitive [N,1,1] or per pixel [N,H,W])."""
    dt = conics.dtype
    yy, xx = torch.meshgrid(torch.arange(height, dtype=dt) + 0.5, torch.arange(width, dtype=dt) + 0.5,
                            indexing="ij")
    a, b, cc = (conics[:, k, None, None] for k in range(3))
    amp = (opac * mus)[:, None, None]
    if beta2d is None:
        dx, dy = xx - mx, yy - my
        sigma = 0.5 * (a * dx * dx + cc * dy * dy) + b * dx * dy
        alpha = amp * torch.exp(-sigma)
        keep = (sigma >= 0) & (alpha.abs() >= 1e-5)
    else:
        dx, dy = mx - xx, my - yy
        d2 = a * dx * dx + 2 * b * dx * dy + cc * dy * dy
        inside = (d2 >= 0) & (d2 < 1)
        h = torch.where(inside, 1 - d2, torch.ones_like(d2))
        alpha = torch.where(inside, amp * h.pow(beta2d[:, None, None]), torch.zeros_like(d2))
        keep = inside & (alpha.abs() >= 1e-5)
    if mask is not None:
        keep = keep & torch.as_tensor(mask)
    if visible is not None:
        keep = keep & visible[:, None, None]
    return torch.where(keep, alpha, torch.zeros_like(alpha))


def rasterize(means2d, conics, opac, mus, width, height, *, beta2d=None, unc=None, detach=False, mask=None,
              visible=None):
    """Dense per-pixel sums (differentiable). Gaussian: delta = pixel - mean,
    sigma = 1/2 delta^T P delta, alpha = o mu e^-sigma, kept when sigma >= 0 and
    |alpha| >= 1e-5. Beta: delta = mean - pixel, d2 = delta^T P delta, alpha =
    o mu (1 - d2)^beta kept for 0 <= d2 < 1 and |alpha| >= 1e-5. variance =
    sum q alpha^2 (alpha detached when `detach`). `mask` [C,N,H,W] restricts
    (pixel, primitive) pairs to intersected tiles."""
    C, N = means2d.shape[:2]
    images, variances = [], []
    for c in range(C):
        alpha = _camera_alpha(means2d[c, :, 0, None, None], means2d[c, :, 1, None, None], conics[c], opac[c], mus[c],
                              width, height, None if beta2d is None else beta2d[c],
                              None if mask is None else mask[c], None if visible is None else visible[c])
        images.append(alpha.sum(0))
        if unc is not None:
            branch = alpha.detach() if detach else alpha
            variances.append((unc[:, None, None] * branch.square()).sum(0))
    image = torch.stack(images)
    return image if unc is None else (image, torch.stack(variances))


def abs_mean2d_grad(means2d, conics, opac, mus, width, height, v, *, beta2d=None, unc=None, v_var=None,
                    detach=False, mask=None):
    """The absgrad output [C,N,2]: sum over pixels of |d(G alpha)/d mean2d| per
    component, G = v (plus 2 q alpha v_var unless detached when a variance
    adjoint is given) held constant per pixel (RasterizeToPixelsBwd.cu:219-221,
    250-255, 455-466). Per-pixel partials come from autograd through
    per-pixel copies of the means."""
    C, N = means2d.shape[:2]
    out = torch.zeros(C, N, 2, dtype=means2d.dtype)
    for c in range(C):
        mx = means2d[c, :, 0, None, None].detach().expand(N, height, width).clone().requires_grad_()
        my = means2d[c, :, 1, None, None].detach().expand(N, height, width).clone().requires_grad_()
        alpha = _camera_alpha(mx, my, conics[c].detach(), opac[c].detach(), mus[c].detach(), width, height,
                              None if beta2d is None else beta2d[c].detach(), None if mask is None else mask[c])
        G = v[c].expand_as(alpha)
        if v_var is not None and unc is not None and not detach:
            G = G + 2 * unc.detach()[:, None, None] * alpha.detach() * v_var[c]
        gx, gy = torch.autograd.grad((G.detach() * alpha).sum(), (mx, my))
        out[c, :, 0] = gx.abs().sum((1, 2))
        out[c, :, 1] = gy.abs().sum((1, 2))
    return out


def rasterize_scene(means, quats, scales, densities, viewmats, Ks, image_size, *, beta=None, uncertainty=None,
                    eps=1e-8, detach=False, mask=None):
    """Physical parameters -> image (or (mean, variance)); differentiable in all
    primitive parameters. Orchestration of r2splat/explicit.py:160-183."""
    height, width = image_size
    cov3 = covariance(quats, scales)
    means2d, cov2d, visible = project(means, cov3, viewmats, Ks, width, height)
    conics, mus = conics_mu(cov2d, eps)
    C = viewmats.shape[0]
    opac = densities[None].expand(C, -1)
    beta2d = None
    if beta is not None:
        mus = mus * beta_ratio(beta)[None]
        beta2d = (beta + 0.5)[None].expand(C, -1)
    return rasterize(means2d, conics, opac, mus, width, height, beta2d=beta2d, unc=uncertainty, detach=detach,
                     mask=mask, visible=visible)


# ---------------------------------------------------------------------------
# voxelization (WorldToVolumeMappingFwd.cu:57-157, VoxelizeToVoxelsFwd.cu:181-219)
# ---------------------------------------------------------------------------
def volume_mapping(means, cov3, extent, center, shape_zyx, eps):
    """Voxel-frame means [N,3], precision P = (r Sigma r + eps I)^-1 [N,3,3] and
    validity (positive definite), r = dims / extent per axis."""
    D, H, W = shape_zyx
    dims = torch.tensor([W, H, D], dtype=means.dtype)
    r = dims / extent
    mapped = (means - center) * r + dims / 2
    Sv = cov3 * r[:, None] * r[None, :] + eps * torch.eye(3, dtype=means.dtype)
    L, info = torch.linalg.cholesky_ex(Sv)
    ok = info == 0
    safe = torch.where(ok[:, None, None], Sv, torch.eye(3, dtype=means.dtype).expand_as(Sv))
    return mapped, torch.linalg.inv(safe), ok


def voxelize(means, quats, scales, densities, shape_zyx, extent, center, *, beta=None, uncertainty=None, eps=1e-8,
             detach=False, mask=None):
    """Samples at voxel centres in world coordinates: voxel (i,j,k) is at
    center + ((index + 1/2) / size - 1/2) * extent per axis. Gaussian profile
    e^(-d2/2) kept when |alpha| >= 1e-6; Beta (1 - d2)^beta kept for 0 <= d2 < 1
    and |alpha| >= 1e-5; eps is added in voxel units squared."""
    D, H, W = shape_zyx
    dt = means.dtype
    extent = torch.as_tensor(extent, dtype=dt)
    center = torch.as_tensor(center, dtype=dt)
    dims = torch.tensor([W, H, D], dtype=dt)
    cov3 = covariance(quats, scales) + torch.diag(eps * (extent / dims).square())
    precision = torch.linalg.inv(cov3)
    axes = [center[k] + ((torch.arange(n, dtype=dt) + 0.5) / n - 0.5) * extent[k] for k, n in enumerate((W, H, D))]
    zz, yy, xx = torch.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    grid = torch.stack([xx, yy, zz], -1)
    mean_out = torch.zeros((D, H, W), dtype=dt)
    var_out = torch.zeros((D, H, W), dtype=dt)
    for n in range(means.shape[0]):
        delta = grid - means[n]
        d2 = torch.einsum("...i,ij,...j->...", delta, precision[n], delta)
        if beta is None:
            alpha = densities[n] * torch.exp(-0.5 * d2)
            keep = (d2 >= 0) & (alpha.abs() >= 1e-6)
        else:
            inside = (d2 >= 0) & (d2 < 1)
            h = torch.where(inside, 1 - d2, torch.ones_like(d2))
            alpha = torch.where(inside, densities[n] * h.pow(beta[n]), torch.zeros_like(d2))
            keep = inside & (alpha.abs() >= 1e-5)
        if mask is not None:
            keep = keep & torch.as_tensor(mask[n])
        alpha = torch.where(keep, alpha, torch.zeros_like(alpha))
        mean_out = mean_out + alpha
        if uncertainty is not None:
            branch = alpha.detach() if detach else alpha
            var_out = var_out + uncertainty[n] * branch.square()
    return mean_out if uncertainty is None else (mean_out, var_out)


# ---------------------------------------------------------------------------
# losses (VolumeLossFwd.cu:39-153, Loss.cpp:115-153)
# ---------------------------------------------------------------------------
def volume_tv(volume: Tensor) -> Tensor:
    """Anisotropic TV: sum of |forward differences| over all edges / edge count."""
    diffs = [volume[:, 1:] - volume[:, :-1], volume[:, :, 1:] - volume[:, :, :-1],
             volume[:, :, :, 1:] - volume[:, :, :, :-1]]
    edges = max(sum(d.numel() for d in diffs), 1)
    return sum(d.abs().sum() for d in diffs) / edges


def volume_tv_iso(volume: Tensor, eps: float = 1e-8) -> Tensor:
    """Isotropic TV (this build's option beside the reference's anisotropic
    loss, which has none): mean over voxels of sqrt(dz^2 + dy^2 + dx^2 + eps),
    forward differences, zero past the last index of each axis."""
    v = volume
    dz = torch.zeros_like(v)
    dy = torch.zeros_like(v)
    dx = torch.zeros_like(v)
    dz[:, :-1] = v[:, 1:] - v[:, :-1]
    dy[:, :, :-1] = v[:, :, 1:] - v[:, :, :-1]
    dx[:, :, :, :-1] = v[:, :, :, 1:] - v[:, :, :, :-1]
    return torch.sqrt(dz * dz + dy * dy + dx * dx + eps).sum() / max(v.numel(), 1)


def relative_error(actual: Tensor, expected: Tensor) -> float:
    """max |a - b| / max |b| (0 when both vanish)."""
    a = actual.detach().double().cpu()
    b = expected.detach().double().cpu()
    scale = b.abs().max().item() if b.numel() else 0.0
    diff = (a - b).abs().max().item() if b.numel() else 0.0
    if scale == 0.0:
        return diff
    return diff / scale

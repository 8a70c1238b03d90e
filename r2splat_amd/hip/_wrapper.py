"""Operator layer: the reference's low-level op names, served by the HIP library.

Each function keeps the name, argument order, return tuple and error class of
the op of the same name in r2splat/cuda/_wrapper.py:19-112 (pybind table
r2splat/cuda/csrc/ext.cpp:8-182, C++ declarations r2splat/cuda/include/Ops.h).
Outputs are allocated here with PyTorch on the inputs' device; the kernels run
on that device's current stream. Shape, dtype, device and contiguity problems
raise RuntimeError, as the reference's TORCH_CHECKs do.
"""

from __future__ import annotations

import math
import struct
from typing import Optional

import torch

from . import _backend as be
from ._backend import call, ptr

Tensor = torch.Tensor


# ---------------------------------------------------------------------------
# checks
# ---------------------------------------------------------------------------
def _check(name: str, t: Tensor, ndim: Optional[int] = None, dtype=None, ref: Optional[Tensor] = None):
    if not isinstance(t, Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if ndim is not None and t.dim() != ndim:
        raise RuntimeError(f"{name} must be {ndim}D, got shape {tuple(t.shape)}")
    if dtype is not None and t.dtype != dtype:
        raise RuntimeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if ref is not None:
        if t.device != ref.device:
            raise RuntimeError(f"{name} must be on the same CUDA device as its inputs")
        if dtype is None and t.dtype != ref.dtype:
            raise RuntimeError(f"{name} must have dtype {ref.dtype}, got {t.dtype}")


def _require(cond: bool, message: str):
    if not cond:
        raise RuntimeError(message)


def _floating(t: Tensor) -> int:
    if t.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"expected float32 or float64 tensors, got {t.dtype}")
    return be.dtype_code(t.dtype)


class _Launch:
    """Device guard + current-stream handle for one op call."""

    def __init__(self, ref: Tensor):
        self.device = ref.device
        self._guard = torch.cuda.device(ref.device)

    def __enter__(self):
        self._guard.__enter__()
        be.lib()
        self.stream = be.stream_of(self.device)
        return self

    def __exit__(self, *exc):
        return self._guard.__exit__(*exc)


# ---------------------------------------------------------------------------
# parameter transform
# ---------------------------------------------------------------------------
def parameter_transform_fwd(quats: Tensor, scales: Tensor):
    """quats [N,4] (wxyz, unnormalised), scales [N,3] -> (covars [N,6], max_scales [N])."""
    _check("quats", quats, 2)
    _check("scales", scales, 2, ref=quats)
    dt = _floating(quats)
    _require(quats.size(1) == 4, "quats must be [N, 4]")
    _require(scales.size(1) == 3 and scales.size(0) == quats.size(0), "scales must be [N, 3]")
    n = quats.size(0)
    covars = quats.new_empty((n, 6))
    max_scales = quats.new_empty((n,))
    with _Launch(quats) as L:
        call("r2s_parameter_transform_fwd", dt, ptr(quats), ptr(scales), ptr(covars), ptr(max_scales), n, L.stream)
    return covars, max_scales


def parameter_transform_bwd(quats: Tensor, scales: Tensor, v_covars: Tensor):
    """VJP to raw quats [N,4] and physical scales [N,3] from v_covars [N,6]."""
    _check("quats", quats, 2)
    _check("scales", scales, 2, ref=quats)
    _check("v_covars", v_covars, 2, ref=quats)
    dt = _floating(quats)
    n = quats.size(0)
    _require(quats.size(1) == 4, "quats must be [N, 4]")
    _require(tuple(scales.shape) == (n, 3), "scales must be [N, 3]")
    _require(tuple(v_covars.shape) == (n, 6), "v_covars must be [N, 6]")
    v_quats = torch.empty_like(quats)
    v_scales = torch.empty_like(scales)
    with _Launch(quats) as L:
        call("r2s_parameter_transform_bwd", dt, ptr(quats), ptr(scales), ptr(v_covars), ptr(v_quats),
             ptr(v_scales), n, L.stream)
    return v_quats, v_scales


# ---------------------------------------------------------------------------
# projection and conics
# ---------------------------------------------------------------------------
def _camera_checks(means, covars, viewmats, Ks):
    _check("means", means, 2)
    _check("covars", covars, 2, ref=means)
    _check("viewmats", viewmats, 3, ref=means)
    _check("Ks", Ks, 3, ref=means)
    _require(means.size(1) == 3, "means must be [N, 3]")
    _require(covars.size(1) == 6 and covars.size(0) == means.size(0), "covars must be [N, 6]")
    _require(viewmats.shape[1:] == (4, 4), "viewmats must be [C, 4, 4]")
    _require(Ks.shape[1:] == (3, 3) and Ks.size(0) == viewmats.size(0), "Ks must be [C, 3, 3]")
    return viewmats.size(0), means.size(0)


def world_to_image_projection_fwd(means, covars, viewmats, Ks, width, height, eps, radius_clip):
    """-> (means2d [C,N,2], cov2d [C,N,6], radii int32 [C,N,2])."""
    C, N = _camera_checks(means, covars, viewmats, Ks)
    dt = _floating(means)
    _require(float(radius_clip) > 0, f"radius_clip must be positive, got {radius_clip}")
    means2d = means.new_empty((C, N, 2))
    cov2d = means.new_empty((C, N, 6))
    radii = torch.zeros((C, N, 2), dtype=torch.int32, device=means.device)
    with _Launch(means) as L:
        call("r2s_world_to_image_projection_fwd", dt, ptr(means), ptr(covars), ptr(viewmats), ptr(Ks),
             int(width), int(height), float(eps), float(radius_clip), ptr(means2d), ptr(cov2d), ptr(radii),
             C, N, L.stream)
    return means2d, cov2d, radii


def world_to_image_projection_bwd(means, covars, viewmats, Ks, width, height, v_means2d, v_cov2d):
    """-> (v_means [N,3], v_covars [N,6]), summed over cameras."""
    C, N = _camera_checks(means, covars, viewmats, Ks)
    dt = _floating(means)
    _check("v_means2d", v_means2d, 3, ref=means)
    _check("v_cov2d_3x3", v_cov2d, 3, ref=means)
    _require(tuple(v_means2d.shape) == (C, N, 2), "v_means2d must be [C, N, 2]")
    _require(tuple(v_cov2d.shape) == (C, N, 6), "v_cov2d_3x3 must be [C, N, 6]")
    v_means = means.new_empty((N, 3))
    v_covars = means.new_empty((N, 6))
    with _Launch(means) as L:
        call("r2s_world_to_image_projection_bwd", dt, ptr(means), ptr(covars), ptr(viewmats), ptr(Ks),
             int(width), int(height), ptr(v_means2d), ptr(v_cov2d), ptr(v_means), ptr(v_covars), C, N,
             L.stream)
    return v_means, v_covars


def conics_mu_fwd(cov2d, eps):
    """cov2d [...,6] -> (conics [...,3], mus [...])."""
    _check("cov2d_3x3", cov2d)
    dt = _floating(cov2d)
    _require(cov2d.dim() >= 1 and cov2d.size(-1) == 6, "cov2d_3x3 must have last dim 6")
    lead = tuple(cov2d.shape[:-1])
    conics = cov2d.new_empty(lead + (3,))
    mus = cov2d.new_empty(lead)
    total = mus.numel()
    with _Launch(cov2d) as L:
        call("r2s_conics_mu_fwd", dt, ptr(cov2d), float(eps), ptr(conics), ptr(mus), total, L.stream)
    return conics, mus


def conics_mu_bwd(cov2d, eps, conics, mus, v_conics, v_mus):
    """-> v_cov2d [...,6]."""
    _check("cov2d_3x3", cov2d)
    dt = _floating(cov2d)
    for name, t in (("conics", conics), ("mus", mus), ("v_conics", v_conics), ("v_mus", v_mus)):
        _check(name, t, ref=cov2d)
    lead = tuple(cov2d.shape[:-1])
    _require(tuple(conics.shape) == lead + (3,) and tuple(v_conics.shape) == lead + (3,), "conics shape mismatch")
    _require(tuple(mus.shape) == lead and tuple(v_mus.shape) == lead, "mus shape mismatch")
    v_cov2d = torch.empty_like(cov2d)
    with _Launch(cov2d) as L:
        call("r2s_conics_mu_bwd", dt, ptr(cov2d), float(eps), ptr(conics), ptr(mus), ptr(v_conics),
             ptr(v_mus), ptr(v_cov2d), mus.numel(), L.stream)
    return v_cov2d


# ---------------------------------------------------------------------------
# volume mapping
# ---------------------------------------------------------------------------
def world_to_volume_mapping_fwd(means, covars, max_scales, ranges, centers, dim2, dim1, dim0, eps3d, radius_clip):
    """-> (means3d_v [N,3], conics [N,6], radii int32 [N,3]); dim2/dim1/dim0 = x/y/z sizes."""
    _check("means", means, 2)
    for name, t in (("covars", covars), ("max_scales", max_scales), ("ranges", ranges), ("centers", centers)):
        _check(name, t, ref=means)
    dt = _floating(means)
    n = means.size(0)
    _require(means.size(1) == 3 and tuple(covars.shape) == (n, 6), "means/covars must be [N,3]/[N,6]")
    _require(ranges.numel() == 3 and centers.numel() == 3, "ranges and centers must have 3 entries")
    mapped = means.new_empty((n, 3))
    conics = means.new_empty((n, 6))
    radii = torch.zeros((n, 3), dtype=torch.int32, device=means.device)
    with _Launch(means) as L:
        call("r2s_world_to_volume_mapping_fwd", dt, ptr(means), ptr(covars), ptr(max_scales), ptr(ranges),
             ptr(centers), int(dim2), int(dim1), int(dim0), float(eps3d), float(radius_clip), ptr(mapped),
             ptr(conics), ptr(radii), n, L.stream)
    return mapped, conics, radii


def world_to_volume_mapping_bwd(means, covars, ranges, centers, dim2, dim1, dim0, eps3d, v_means3d_v, v_conics):
    """-> (v_means [N,3], v_covars [N,6])."""
    _check("means", means, 2)
    for name, t in (("covars", covars), ("ranges", ranges), ("centers", centers),
                    ("v_means3d_v", v_means3d_v), ("v_conics", v_conics)):
        _check(name, t, ref=means)
    dt = _floating(means)
    n = means.size(0)
    _require(tuple(v_means3d_v.shape) == (n, 3) and tuple(v_conics.shape) == (n, 6), "adjoint shape mismatch")
    v_means = means.new_empty((n, 3))
    v_covars = means.new_empty((n, 6))
    with _Launch(means) as L:
        call("r2s_world_to_volume_mapping_bwd", dt, ptr(means), ptr(covars), ptr(ranges), ptr(centers),
             int(dim2), int(dim1), int(dim0), float(eps3d), ptr(v_means3d_v), ptr(v_conics), ptr(v_means),
             ptr(v_covars), n, L.stream)
    return v_means, v_covars


# ---------------------------------------------------------------------------
# tile binning
# ---------------------------------------------------------------------------
def _workspace(nbytes: int, device) -> Tensor:
    return torch.empty((max(int(nbytes), 1),), dtype=torch.uint8, device=device)


def isect_tiles(means2d, radii, tile_size, n_tiles_dim1, n_tiles_dim0):
    """-> (tiles_per_gauss int32 [C,N], isect_ids int64 [I], flatten_ids int32 [I]).

    Keys are cam * n_tiles + ty * n_tiles_x + tx and are sorted stably, so each
    tile's list is in ascending cam * N + g (IsectTiles.cpp:31-170).
    """
    _check("means2d", means2d, 3)
    _check("radii", radii, dtype=torch.int32)
    _require(radii.device == means2d.device, "radii must be on the same CUDA device as means2d")
    dt = _floating(means2d)
    _require(means2d.size(2) == 2, "means2d last dim must be 2")
    _require(radii.dim() == 2 or (radii.dim() == 3 and radii.size(2) == 2), "radii must be [C,N] or [C,N,2]")
    C, N = means2d.size(0), means2d.size(1)
    _require(radii.size(0) == C and radii.size(1) == N, "radii must match means2d in C and N")
    rect = 1 if radii.dim() == 3 else 0
    dev = means2d.device
    tiles = torch.empty((C, N), dtype=torch.int32, device=dev)
    total = C * N
    cum = torch.empty((total + 1,), dtype=torch.int32, device=dev)
    d_total = torch.empty((1,), dtype=torch.int32, device=dev)
    with _Launch(means2d) as L:
        scan_ws = _workspace(be.lib().r2s_scan_workspace_bytes(total), dev)
        call("r2s_isect_tiles_count", dt, ptr(means2d), ptr(radii), rect, C, N, int(tile_size),
             int(n_tiles_dim1), int(n_tiles_dim0), ptr(tiles), ptr(cum), ptr(scan_ws), ptr(d_total), L.stream)
        n_isects = int(d_total.item())
        isect_ids = torch.empty((n_isects,), dtype=torch.int64, device=dev)
        flatten_ids = torch.empty((n_isects,), dtype=torch.int32, device=dev)
        if n_isects > 0:
            max_key = C * int(n_tiles_dim1) * int(n_tiles_dim0)
            end_bit = math.ceil(math.log2(max_key + 1))
            nbytes = be.lib().r2s_sort_workspace_bytes(n_isects, end_bit)
            sort_ws = _workspace(nbytes, dev)
            call("r2s_isect_tiles_emit", dt, ptr(means2d), ptr(radii), rect, C, N, int(tile_size),
                 int(n_tiles_dim1), int(n_tiles_dim0), ptr(cum), n_isects, ptr(isect_ids), ptr(flatten_ids),
                 ptr(sort_ws), nbytes, L.stream)
    return tiles, isect_ids, flatten_ids


def _offsets(isect_ids, n_keys, shape):
    _check("isect_ids", isect_ids, 1, dtype=torch.int64)
    offsets = torch.empty(shape, dtype=torch.int32, device=isect_ids.device)
    with _Launch(isect_ids) as L:
        call("r2s_isect_offset_encode", ptr(isect_ids), isect_ids.numel(), int(n_keys), ptr(offsets), L.stream)
    return offsets


def isect_offset_encode(isect_ids, C, n_tiles_dim1, n_tiles_dim0):
    """-> tile_offsets int32 [C, n_tiles_y, n_tiles_x]."""
    C, tx, ty = int(C), int(n_tiles_dim1), int(n_tiles_dim0)
    return _offsets(isect_ids, C * tx * ty, (C, ty, tx))


def isect_tiles_3d(means3d, radii, tile_size, n_tiles_dim2, n_tiles_dim1, n_tiles_dim0):
    """-> (tiles_per_gauss int32 [N], isect_ids int64 [I], flatten_ids int32 [I]); x/y/z tile counts."""
    return _isect_tiles_3d(means3d, radii, tile_size, n_tiles_dim2, n_tiles_dim1, n_tiles_dim0)[:3]


def _isect_tiles_3d(means3d, radii, tile_size, n_tiles_dim2, n_tiles_dim1, n_tiles_dim0):
    """isect_tiles_3d plus the exclusive scan of the tile counts, int32 [N + 1]
    (the first emission index of each primitive)."""
    _check("means3d", means3d, 2)
    _check("radii", radii, 2, dtype=torch.int32)
    _require(radii.device == means3d.device, "radii must be on the same CUDA device as means3d")
    dt = _floating(means3d)
    _require(means3d.size(1) == 3 and radii.size(1) == 3, "means3d and radii must be [N, 3]")
    N = means3d.size(0)
    _require(radii.size(0) == N, "N dimension must match")
    nx, ny, nz = int(n_tiles_dim2), int(n_tiles_dim1), int(n_tiles_dim0)
    dev = means3d.device
    tiles = torch.empty((N,), dtype=torch.int32, device=dev)
    cum = torch.empty((N + 1,), dtype=torch.int32, device=dev)
    d_total = torch.empty((1,), dtype=torch.int32, device=dev)
    with _Launch(means3d) as L:
        scan_ws = _workspace(be.lib().r2s_scan_workspace_bytes(N), dev)
        call("r2s_isect_tiles_3d_count", dt, ptr(means3d), ptr(radii), N, int(tile_size), nx, ny, nz,
             ptr(tiles), ptr(cum), ptr(scan_ws), ptr(d_total), L.stream)
        n_isects = int(d_total.item())
        isect_ids = torch.empty((n_isects,), dtype=torch.int64, device=dev)
        flatten_ids = torch.empty((n_isects,), dtype=torch.int32, device=dev)
        if n_isects > 0:
            end_bit = math.ceil(math.log2(nx * ny * nz + 1))
            nbytes = be.lib().r2s_sort_workspace_bytes(n_isects, end_bit)
            sort_ws = _workspace(nbytes, dev)
            call("r2s_isect_tiles_3d_emit", dt, ptr(means3d), ptr(radii), N, int(tile_size), nx, ny, nz,
                 ptr(cum), n_isects, ptr(isect_ids), ptr(flatten_ids), ptr(sort_ws), nbytes, L.stream)
    return tiles, isect_ids, flatten_ids, cum


def isect_offset_encode_3d(isect_ids, n_tiles_dim2, n_tiles_dim1, n_tiles_dim0):
    """-> tile_offsets int32 [n_tiles_z, n_tiles_y, n_tiles_x]."""
    nx, ny, nz = int(n_tiles_dim2), int(n_tiles_dim1), int(n_tiles_dim0)
    return _offsets(isect_ids, nx * ny * nz, (nz, ny, nx))


# ---------------------------------------------------------------------------
# 2D rasterization
# ---------------------------------------------------------------------------
def _raster_inputs(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids):
    _check("means2d", means2d, 3)
    dt = _floating(means2d)
    _check("conics", conics, 3, ref=means2d)
    _check("opacities", opacities, 2, ref=means2d)
    _check("mus", mus, 2, ref=means2d)
    if betas is not None:
        _check("betas", betas, 2, ref=means2d)
    _check("tile_offsets", tile_offsets, 3, dtype=torch.int32)
    _check("flatten_ids", flatten_ids, 1, dtype=torch.int32)
    C, N = means2d.size(0), means2d.size(1)
    _require(means2d.size(2) == 2 and tuple(conics.shape) == (C, N, 3), "conics shape must match means2d [C, N, 3]")
    _require(tuple(opacities.shape) == (C, N) and tuple(mus.shape) == (C, N), "opacities/mus must be [C, N]")
    _require(betas is None or tuple(betas.shape) == (C, N), "betas must be [C, N]")
    _require(tile_offsets.size(0) == C, "tile_offsets C dimension must match")
    return dt, C, N, tile_offsets.size(2), tile_offsets.size(1)


def _image_adjoint(name, t, ref, C, H, W):
    _check(name, t, 3, ref=ref)
    _require(tuple(t.shape) == (C, H, W), f"{name} shape must be [C, dim0, dim1]")


def _raster_fwd(means2d, conics, opacities, mus, betas, uncertainties, tile_offsets, flatten_ids, width, height, tile_size):
    dt, C, N, ntx, nty = _raster_inputs(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids)
    W, H, T = int(width), int(height), int(tile_size)
    if uncertainties is not None:
        _check("uncertainties", uncertainties, 1, ref=means2d)
        _require(uncertainties.numel() == N, "uncertainties must be [N]")
    out = means2d.new_zeros((C, H, W))
    var = means2d.new_zeros((C, H, W)) if uncertainties is not None else None
    with _Launch(means2d) as L:
        common = (ptr(tile_offsets), ptr(flatten_ids), flatten_ids.numel(), C, N, W, H, T, ntx, nty)
        if betas is None and var is None:
            call("r2s_rasterize_to_pixels_fwd", dt, ptr(means2d), ptr(conics), ptr(opacities), ptr(mus), *common,
                 ptr(out), L.stream)
        elif betas is None:
            call("r2s_rasterize_to_pixels_mean_var_fwd", dt, ptr(means2d), ptr(conics), ptr(opacities), ptr(mus),
                 ptr(uncertainties), *common, ptr(out), ptr(var), L.stream)
        elif var is None:
            call("r2s_beta_rasterize_to_pixels_fwd", dt, ptr(means2d), ptr(conics), ptr(opacities), ptr(mus),
                 ptr(betas), *common, ptr(out), L.stream)
        else:
            call("r2s_beta_rasterize_to_pixels_mean_var_fwd", dt, ptr(means2d), ptr(conics), ptr(opacities),
                 ptr(mus), ptr(betas), ptr(uncertainties), *common, ptr(out), ptr(var), L.stream)
    return out if var is None else (out, var)


def rasterize_to_pixels_fwd(means2d, conics, opacities, mus, tile_offsets, flatten_ids, dim1, dim0, tile_size):
    """-> image [C, dim0, dim1]."""
    return _raster_fwd(means2d, conics, opacities, mus, None, None, tile_offsets, flatten_ids, dim1, dim0, tile_size)


def rasterize_to_pixels_mean_var_fwd(means2d, conics, opacities, mus, uncertainties, tile_offsets, flatten_ids,
                                     dim1, dim0, tile_size):
    """-> (image, variance) [C, dim0, dim1]."""
    return _raster_fwd(means2d, conics, opacities, mus, None, uncertainties, tile_offsets, flatten_ids,
                       dim1, dim0, tile_size)


def beta_rasterize_to_pixels_fwd(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids, dim1, dim0,
                                 tile_size):
    return _raster_fwd(means2d, conics, opacities, mus, betas, None, tile_offsets, flatten_ids, dim1, dim0, tile_size)


def beta_rasterize_to_pixels_mean_var_fwd(means2d, conics, opacities, mus, betas, uncertainties, tile_offsets,
                                          flatten_ids, dim1, dim0, tile_size):
    return _raster_fwd(means2d, conics, opacities, mus, betas, uncertainties, tile_offsets, flatten_ids,
                       dim1, dim0, tile_size)


def _uncertainty_bwd(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids, v_render_var, width,
                     height, tile_size):
    dt, C, N, ntx, nty = _raster_inputs(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids)
    W, H = int(width), int(height)
    _image_adjoint("v_render_var", v_render_var, means2d, C, H, W)
    v_q = means2d.new_empty((N,))
    with _Launch(means2d) as L:
        call("r2s_rasterize_to_pixels_uncertainty_bwd", dt, ptr(means2d), ptr(conics), ptr(opacities), ptr(mus),
             ptr(betas), ptr(tile_offsets), ptr(flatten_ids), flatten_ids.numel(), C, N, W, H, int(tile_size),
             ntx, nty, ptr(v_render_var), ptr(v_q), L.stream)
    return v_q


def rasterize_to_pixels_uncertainty_bwd(means2d, conics, opacities, mus, tile_offsets, flatten_ids, v_render_var,
                                        dim1, dim0, tile_size):
    """-> v_uncertainties [N]."""
    return _uncertainty_bwd(means2d, conics, opacities, mus, None, tile_offsets, flatten_ids, v_render_var,
                            dim1, dim0, tile_size)


def beta_rasterize_to_pixels_uncertainty_bwd(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids,
                                             v_render_var, dim1, dim0, tile_size):
    return _uncertainty_bwd(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids, v_render_var,
                            dim1, dim0, tile_size)


def _raster_grads(means2d, betas, absgrad):
    v_abs = torch.empty_like(means2d) if absgrad else None
    out = [v_abs, torch.empty_like(means2d), means2d.new_empty(means2d.shape[:2] + (3,)),
           means2d.new_empty(means2d.shape[:2]), means2d.new_empty(means2d.shape[:2])]
    if betas is not None:
        out.append(means2d.new_empty(means2d.shape[:2]))
    return out


def _raster_bwd_pixel(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids, v_render, width, height,
                      tile_size, absgrad):
    dt, C, N, ntx, nty = _raster_inputs(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids)
    W, H = int(width), int(height)
    _image_adjoint("v_render_output", v_render, means2d, C, H, W)
    g = _raster_grads(means2d, betas, absgrad)
    with _Launch(means2d) as L:
        call("r2s_rasterize_to_pixels_bwd", dt, ptr(means2d), ptr(conics), ptr(opacities), ptr(mus), ptr(betas),
             ptr(tile_offsets), ptr(flatten_ids), flatten_ids.numel(), C, N, W, H, int(tile_size), ntx, nty,
             ptr(v_render), ptr(g[0]), ptr(g[1]), ptr(g[2]), ptr(g[3]), ptr(g[4]),
             ptr(g[5]) if betas is not None else None, L.stream)
    return tuple(g)


def _raster_bwd_isect(means2d, conics, opacities, mus, betas, tile_offsets, isect_ids, flatten_ids, v_render,
                      v_render_var, uncertainties, detach, width, height, tile_size, absgrad):
    dt, C, N, ntx, nty = _raster_inputs(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids)
    W, H = int(width), int(height)
    _check("isect_ids", isect_ids, 1, dtype=torch.int64)
    _require(isect_ids.numel() == flatten_ids.numel(), "isect_ids and flatten_ids must have the same length")
    _image_adjoint("v_render_output", v_render, means2d, C, H, W)
    g = _raster_grads(means2d, betas, absgrad)
    v_q = None
    unc = None
    if v_render_var is not None:
        _image_adjoint("v_render_var", v_render_var, means2d, C, H, W)
        v_q = means2d.new_empty((N,))
        if not detach:
            _check("uncertainties", uncertainties, 1, ref=means2d)
            _require(uncertainties.numel() == N, "uncertainties shape must match [N]")
            unc = uncertainties
    with _Launch(means2d) as L:
        call("r2s_rasterize_to_pixels_bwd_per_gaussian", dt, ptr(means2d), ptr(conics), ptr(opacities), ptr(mus),
             ptr(betas), ptr(isect_ids), ptr(flatten_ids), flatten_ids.numel(), C, N, W, H, int(tile_size), ntx,
             nty, ptr(v_render), ptr(v_render_var), ptr(unc), ptr(g[0]), ptr(g[1]), ptr(g[2]), ptr(g[3]),
             ptr(g[4]), ptr(g[5]) if betas is not None else None, ptr(v_q), L.stream)
    if v_q is not None:
        g.append(v_q)
    return tuple(g)


def rasterize_to_pixels_bwd(means2d, conics, opacities, mus, tile_offsets, flatten_ids, v_render_output, dim1, dim0,
                            tile_size, absgrad):
    """Pixel-mode VJP -> (v_means2d_abs | None, v_means2d, v_conics, v_opacities, v_mus)."""
    return _raster_bwd_pixel(means2d, conics, opacities, mus, None, tile_offsets, flatten_ids, v_render_output,
                             dim1, dim0, tile_size, absgrad)


def beta_rasterize_to_pixels_bwd(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids, v_render_output,
                                 dim1, dim0, tile_size, absgrad):
    """-> (v_abs | None, v_means2d, v_conics, v_opacities, v_mus, v_betas)."""
    return _raster_bwd_pixel(means2d, conics, opacities, mus, betas, tile_offsets, flatten_ids, v_render_output,
                             dim1, dim0, tile_size, absgrad)


def rasterize_to_pixels_bwd_per_gaussian(means2d, conics, opacities, mus, tile_offsets, isect_ids, flatten_ids,
                                         v_render_output, dim1, dim0, tile_size, absgrad):
    """Primitive-mode VJP -> (v_abs | None, v_means2d, v_conics, v_opacities, v_mus)."""
    return _raster_bwd_isect(means2d, conics, opacities, mus, None, tile_offsets, isect_ids, flatten_ids,
                             v_render_output, None, None, True, dim1, dim0, tile_size, absgrad)


def beta_rasterize_to_pixels_bwd_per_gaussian(means2d, conics, opacities, mus, betas, tile_offsets, isect_ids,
                                              flatten_ids, v_render_output, dim1, dim0, tile_size, absgrad):
    return _raster_bwd_isect(means2d, conics, opacities, mus, betas, tile_offsets, isect_ids, flatten_ids,
                             v_render_output, None, None, True, dim1, dim0, tile_size, absgrad)


def _split_mean_var(args, n_fixed):
    """Resolves the two mean_var VJP overloads (detached form without
    (uncertainties, detach)) as in r2splat/cuda/csrc/ext.cpp:106-128."""
    fixed, rest = list(args[:n_fixed]), list(args[n_fixed:])
    if len(rest) == 4:
        return fixed, None, True, rest
    if len(rest) == 6:
        return fixed, rest[0], bool(rest[1]), rest[2:]
    raise TypeError(f"unexpected argument count {len(args)}")


def rasterize_to_pixels_mean_var_bwd_per_gaussian(*args):
    """(means2d, conics, opacities, mus, tile_offsets, isect_ids, flatten_ids, v_render, v_render_var,
    [uncertainties, detach,] dim1, dim0, tile_size, absgrad)
    -> (v_abs | None, v_means2d, v_conics, v_opacities, v_mus, v_uncertainties)."""
    fixed, unc, detach, (w, h, t, absgrad) = _split_mean_var(args, 9)
    m, c, o, mu, off, ids, flat, v, vv = fixed
    return _raster_bwd_isect(m, c, o, mu, None, off, ids, flat, v, vv, unc, detach, w, h, t, absgrad)


def beta_rasterize_to_pixels_mean_var_bwd_per_gaussian(*args):
    """Beta form: betas follows mus; returns (..., v_mus, v_betas, v_uncertainties)."""
    fixed, unc, detach, (w, h, t, absgrad) = _split_mean_var(args, 10)
    m, c, o, mu, b, off, ids, flat, v, vv = fixed
    return _raster_bwd_isect(m, c, o, mu, b, off, ids, flat, v, vv, unc, detach, w, h, t, absgrad)


# ---------------------------------------------------------------------------
# 3D voxelization
# ---------------------------------------------------------------------------
def _vox_inputs(means3d, conics, opacities, betas, tile_offsets, flatten_ids):
    _check("means3d", means3d, 2)
    dt = _floating(means3d)
    _check("conics", conics, 2, ref=means3d)
    _check("opacities", opacities, 1, ref=means3d)
    if betas is not None:
        _check("betas", betas, 1, ref=means3d)
    _check("tile_offsets", tile_offsets, 3, dtype=torch.int32)
    _check("flatten_ids", flatten_ids, 1, dtype=torch.int32)
    N = means3d.size(0)
    _require(means3d.size(1) == 3 and tuple(conics.shape) == (N, 6), "means3d/conics must be [N,3]/[N,6]")
    _require(opacities.numel() == N and (betas is None or betas.numel() == N), "opacities/betas must be [N]")
    ntz, nty, ntx = tile_offsets.shape
    return dt, N, ntx, nty, ntz


def _vox_fwd(means3d, conics, opacities, betas, unc, tile_offsets, flatten_ids, dim2, dim1, dim0, tile_size):
    dt, N, ntx, nty, ntz = _vox_inputs(means3d, conics, opacities, betas, tile_offsets, flatten_ids)
    dx, dy, dz = int(dim2), int(dim1), int(dim0)
    if unc is not None:
        _check("uncertainties", unc, 1, ref=means3d)
        _require(unc.numel() == N, "uncertainties must be [N]")
    out = means3d.new_zeros((dz, dy, dx))
    var = means3d.new_zeros((dz, dy, dx)) if unc is not None else None
    with _Launch(means3d) as L:
        call("r2s_voxelize_to_voxels_fwd", dt, ptr(means3d), ptr(conics), ptr(opacities), ptr(betas), ptr(unc),
             ptr(tile_offsets), ptr(flatten_ids), flatten_ids.numel(), N, dx, dy, dz, int(tile_size), ntx, nty, ntz,
             ptr(out), ptr(var), L.stream)
    return out if var is None else (out, var)


def voxelize_to_voxels_fwd(means3d, conics, opacities, tile_offsets, flatten_ids, dim2, dim1, dim0, tile_size):
    """-> volume [dim0, dim1, dim2] (z, y, x)."""
    return _vox_fwd(means3d, conics, opacities, None, None, tile_offsets, flatten_ids, dim2, dim1, dim0, tile_size)


def voxelize_to_voxels_mean_var_fwd(means3d, conics, opacities, uncertainties, tile_offsets, flatten_ids, dim2,
                                    dim1, dim0, tile_size):
    return _vox_fwd(means3d, conics, opacities, None, uncertainties, tile_offsets, flatten_ids, dim2, dim1, dim0,
                    tile_size)


def beta_voxelize_to_voxels_fwd(means3d, conics, opacities, betas, tile_offsets, flatten_ids, dim2, dim1, dim0,
                                tile_size):
    return _vox_fwd(means3d, conics, opacities, betas, None, tile_offsets, flatten_ids, dim2, dim1, dim0, tile_size)


def beta_voxelize_to_voxels_mean_var_fwd(means3d, conics, opacities, betas, uncertainties, tile_offsets,
                                         flatten_ids, dim2, dim1, dim0, tile_size):
    return _vox_fwd(means3d, conics, opacities, betas, uncertainties, tile_offsets, flatten_ids, dim2, dim1, dim0,
                    tile_size)


def _vox_grads(means3d, betas):
    N = means3d.size(0)
    g = [torch.empty_like(means3d), means3d.new_empty((N, 6)), means3d.new_empty((N,))]
    if betas is not None:
        g.append(means3d.new_empty((N,)))
    return g


def _vox_volume_adjoint(name, t, ref, dx, dy, dz):
    _check(name, t, 3, ref=ref)
    _require(tuple(t.shape) == (dz, dy, dx), f"{name} shape must be [dim0, dim1, dim2]")


def _vox_bwd_tile(means3d, conics, opacities, betas, tile_offsets, flatten_ids, v_render, dim2, dim1, dim0,
                  tile_size):
    dt, N, ntx, nty, ntz = _vox_inputs(means3d, conics, opacities, betas, tile_offsets, flatten_ids)
    dx, dy, dz = int(dim2), int(dim1), int(dim0)
    _vox_volume_adjoint("v_render_output", v_render, means3d, dx, dy, dz)
    g = _vox_grads(means3d, betas)
    with _Launch(means3d) as L:
        call("r2s_voxelize_to_voxels_bwd", dt, ptr(means3d), ptr(conics), ptr(opacities), ptr(betas),
             ptr(tile_offsets), ptr(flatten_ids), flatten_ids.numel(), N, dx, dy, dz, int(tile_size), ntx, nty, ntz,
             ptr(v_render), ptr(g[0]), ptr(g[1]), ptr(g[2]), ptr(g[3]) if betas is not None else None, L.stream)
    return tuple(g)


def _vox_bwd_isect(means3d, conics, opacities, betas, tile_offsets, isect_ids, flatten_ids, v_render, v_var, unc,
                   detach, dim2, dim1, dim0, tile_size):
    dt, N, ntx, nty, ntz = _vox_inputs(means3d, conics, opacities, betas, tile_offsets, flatten_ids)
    dx, dy, dz = int(dim2), int(dim1), int(dim0)
    _check("isect_ids", isect_ids, 1, dtype=torch.int64)
    _require(isect_ids.numel() == flatten_ids.numel(), "isect_ids and flatten_ids must have the same length")
    _vox_volume_adjoint("v_render_output", v_render, means3d, dx, dy, dz)
    g = _vox_grads(means3d, betas)
    v_q = None
    u = None
    if v_var is not None:
        _vox_volume_adjoint("v_render_var", v_var, means3d, dx, dy, dz)
        v_q = means3d.new_empty((N,))
        if not detach:
            _check("uncertainties", unc, 1, ref=means3d)
            _require(unc.numel() == N, "uncertainties shape must match [N]")
            u = unc
    with _Launch(means3d) as L:
        call("r2s_voxelize_to_voxels_bwd_per_gaussian", dt, ptr(means3d), ptr(conics), ptr(opacities), ptr(betas),
             ptr(isect_ids), ptr(flatten_ids), flatten_ids.numel(), N, dx, dy, dz, int(tile_size), ntx, nty, ntz,
             ptr(v_render), ptr(v_var), ptr(u), ptr(g[0]), ptr(g[1]), ptr(g[2]),
             ptr(g[3]) if betas is not None else None, ptr(v_q), L.stream)
    if v_q is not None:
        g.append(v_q)
    return tuple(g)


def voxelize_to_voxels_bwd(means3d, conics, opacities, tile_offsets, flatten_ids, v_render_output, dim2, dim1, dim0,
                           tile_size):
    """Tile-parallel VJP -> (v_means3d, v_conics, v_opacities)."""
    return _vox_bwd_tile(means3d, conics, opacities, None, tile_offsets, flatten_ids, v_render_output, dim2, dim1,
                         dim0, tile_size)


def beta_voxelize_to_voxels_bwd(means3d, conics, opacities, betas, tile_offsets, flatten_ids, v_render_output, dim2,
                                dim1, dim0, tile_size):
    return _vox_bwd_tile(means3d, conics, opacities, betas, tile_offsets, flatten_ids, v_render_output, dim2, dim1,
                         dim0, tile_size)


def voxelize_to_voxels_bwd_per_gaussian(means3d, conics, opacities, tile_offsets, isect_ids, flatten_ids,
                                        v_render_output, dim2, dim1, dim0, tile_size):
    return _vox_bwd_isect(means3d, conics, opacities, None, tile_offsets, isect_ids, flatten_ids, v_render_output,
                          None, None, True, dim2, dim1, dim0, tile_size)


def beta_voxelize_to_voxels_bwd_per_gaussian(means3d, conics, opacities, betas, tile_offsets, isect_ids, flatten_ids,
                                             v_render_output, dim2, dim1, dim0, tile_size):
    return _vox_bwd_isect(means3d, conics, opacities, betas, tile_offsets, isect_ids, flatten_ids, v_render_output,
                          None, None, True, dim2, dim1, dim0, tile_size)


def _split_vox_mean_var(args, n_fixed):
    fixed, rest = list(args[:n_fixed]), list(args[n_fixed:])
    if len(rest) == 4:
        return fixed, None, True, rest
    if len(rest) == 6:
        return fixed, rest[0], bool(rest[1]), rest[2:]
    raise TypeError(f"unexpected argument count {len(args)}")


def voxelize_to_voxels_mean_var_bwd_per_gaussian(*args):
    """(means3d, conics, opacities, tile_offsets, isect_ids, flatten_ids, v_render, v_render_var,
    [uncertainties, detach,] dim2, dim1, dim0, tile_size) -> (v_means3d, v_conics, v_opacities, v_uncertainties)."""
    fixed, unc, detach, (d2, d1, d0, t) = _split_vox_mean_var(args, 8)
    m, c, o, off, ids, flat, v, vv = fixed
    return _vox_bwd_isect(m, c, o, None, off, ids, flat, v, vv, unc, detach, d2, d1, d0, t)


def beta_voxelize_to_voxels_mean_var_bwd_per_gaussian(*args):
    fixed, unc, detach, (d2, d1, d0, t) = _split_vox_mean_var(args, 9)
    m, c, o, b, off, ids, flat, v, vv = fixed
    return _vox_bwd_isect(m, c, o, b, off, ids, flat, v, vv, unc, detach, d2, d1, d0, t)


# ---------------------------------------------------------------------------
# losses
# ---------------------------------------------------------------------------
def _f32(x) -> float:
    """x rounded to float32: the reference's loss ops take float scalars (Ops.h:14-44)."""
    return struct.unpack("f", struct.pack("f", float(x)))[0]


def volume_loss_fwd(volume, lambda_tv):
    """Anisotropic TV over (B,D,H,W) -> (lambda * tv, tv, tv_map)."""
    _check("volume", volume, 4)
    dt = _floating(volume)
    B, Dd, H, W = volume.shape
    tv_map = torch.empty_like(volume)
    tv_sum = volume.new_empty((1,))
    with _Launch(volume) as L:
        call("r2s_volume_tv_fwd", dt, ptr(volume), B, Dd, H, W, ptr(tv_map), ptr(tv_sum), L.stream)
    edges = max(B * ((Dd - 1) * H * W + Dd * (H - 1) * W + Dd * H * (W - 1)), 1)
    # the reference divides by float(edge_count) and scales by a float lambda (Loss.cpp:115-153)
    tv = (tv_sum / _f32(edges)).squeeze()
    return _f32(lambda_tv) * tv, tv, tv_map


def volume_loss_bwd(volume, grad_output, lambda_tv):
    """-> lambda * grad_output * d tv / d volume."""
    _check("volume", volume, 4)
    _check("grad_output", grad_output)
    _require(grad_output.device == volume.device, "grad_output must be on the same CUDA device as volume")
    dt = _floating(volume)
    _require(grad_output.dtype == volume.dtype and grad_output.numel() >= 1, "grad_output must be a scalar tensor")
    B, Dd, H, W = volume.shape
    grad = torch.empty_like(volume)
    with _Launch(volume) as L:
        call("r2s_volume_tv_bwd", dt, ptr(volume), ptr(grad_output), B, Dd, H, W, _f32(lambda_tv), ptr(grad),
             L.stream)
    return grad


def volume_loss_iso_fwd(volume, lambda_tv, eps=1e-8):
    """Isotropic TV over (B,D,H,W): tv = mean over voxels of sqrt(dz^2 + dy^2 +
    dx^2 + eps) (forward differences, zero past the last index) -> (lambda * tv,
    tv, tv_map). An option beside the reference's anisotropic volume_loss_fwd."""
    _check("volume", volume, 4)
    dt = _floating(volume)
    _require(float(eps) > 0.0, "eps must be positive")
    B, Dd, H, W = volume.shape
    tv_map = torch.empty_like(volume)
    tv_sum = volume.new_empty((1,))
    with _Launch(volume) as L:
        call("r2s_volume_tv_iso_fwd", dt, ptr(volume), B, Dd, H, W, float(eps), ptr(tv_map), ptr(tv_sum), L.stream)
    tv = (tv_sum / max(volume.numel(), 1)).squeeze()
    return (float(lambda_tv) * tv), tv, tv_map


def volume_loss_iso_bwd(volume, grad_output, lambda_tv, eps=1e-8):
    """-> lambda * grad_output * d tv / d volume of volume_loss_iso_fwd."""
    _check("volume", volume, 4)
    _check("grad_output", grad_output)
    _require(grad_output.device == volume.device, "grad_output must be on the same CUDA device as volume")
    dt = _floating(volume)
    _require(grad_output.dtype == volume.dtype and grad_output.numel() >= 1, "grad_output must be a scalar tensor")
    _require(float(eps) > 0.0, "eps must be positive")
    B, Dd, H, W = volume.shape
    grad = torch.empty_like(volume)
    with _Launch(volume) as L:
        call("r2s_volume_tv_iso_bwd", dt, ptr(volume), ptr(grad_output), B, Dd, H, W, float(eps), float(lambda_tv),
             ptr(grad), L.stream)
    return grad


def image_loss_l1_fwd(prediction, target, lambda_l1):
    """Mean L1 over (B,H,W) -> (lambda * l1, l1, l1_map)."""
    _check("prediction", prediction, 3)
    _check("target", target, 3)
    _require(target.device == prediction.device, "target must be on the same CUDA device as prediction")
    _require(prediction.shape == target.shape, "prediction and target must have same shape")
    dt = _floating(prediction)
    l1_map = torch.empty_like(prediction)
    l1_sum = prediction.new_empty((1,))
    with _Launch(prediction) as L:
        call("r2s_image_l1_fwd", dt, ptr(prediction), ptr(target), prediction.numel(), ptr(l1_map), ptr(l1_sum),
             L.stream)
    l1 = (l1_sum / max(prediction.numel(), 1)).squeeze()
    return _f32(lambda_l1) * l1, l1, l1_map


def image_loss_l1_bwd(prediction, target, grad_output, lambda_l1):
    _check("prediction", prediction, 3)
    _check("target", target, 3)
    _check("grad_output", grad_output)
    _require(target.device == prediction.device, "target must be on the same CUDA device as prediction")
    _require(grad_output.device == prediction.device, "grad_output must be on the same CUDA device as prediction")
    _require(prediction.shape == target.shape, "prediction and target must have same shape")
    dt = _floating(prediction)
    grad = torch.empty_like(prediction)
    with _Launch(prediction) as L:
        call("r2s_image_l1_bwd", dt, ptr(prediction), ptr(target), ptr(grad_output), prediction.numel(),
             _f32(lambda_l1), ptr(grad), L.stream)
    return grad


image_loss_fwd = image_loss_l1_fwd
image_loss_bwd = image_loss_l1_bwd


def is_cuda_extension_available() -> bool:
    return be.is_available()


__all__ = [name for name in dir() if not name.startswith("_") and name not in ("annotations", "math", "struct", "torch",
                                                                                   "Tensor", "Optional", "be", "call",
                                                                                   "ptr")]

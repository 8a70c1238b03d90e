"""Explicit forward / VJP calls for physical primitive parameters.

Drop-in for r2splat/explicit.py:102-351: ``rasterize_fwd`` / ``voxelize_fwd``
return ``(output, ctx)`` without building an autograd graph, and
``rasterize_bwd`` / ``voxelize_bwd`` consume the context once and return
``ParameterGradients`` without touching ``.grad``. Geometry (cameras, volume
frame) is constant; activations and loss derivatives belong to the caller.

Each stage is one HIP operator from ``r2splat_amd.hip`` (the reference's op
sequence: parameter transform -> projection -> conics/mu -> tile binning ->
rasterization, and the VJPs in reverse).
"""

from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional, Sequence, Union

import torch

from . import _validation as check
from . import fused
from . import hip as ops
from .hip import _backend as be
from .hip import _wrapper as _ops_impl
from ._beta import beta_path_log_grad, beta_path_ratio
from .profiling import stage

Tensor = torch.Tensor
Output = Union[Tensor, tuple]

RASTER_TILE = 16
VOXEL_TILE = 8
_PRIMITIVE_NAMES = ("means", "quats", "scales", "densities", "beta", "uncertainty")

# "fused": float32 projections (Gaussian or Beta, with or without
# uncertainty) run the fused pipeline (r2splat_amd.fused); every other case,
# and "ops", runs one operator per stage.
BACKEND = os.environ.get("R2S_BACKEND", "fused")


def _use_fused(means) -> bool:
    return BACKEND == "fused" and means.dtype == torch.float32


def _use_fused_voxels(means) -> bool:
    """float32 voxel forwards and VJPs (mean or (mean, variance), 8^3 tiles) run
    the fused kernels (fused.voxel_render / fused.voxel_vjp)."""
    return BACKEND == "fused" and means.dtype == torch.float32 and VOXEL_TILE == 8


@dataclass
class ParameterGradients:
    """VJPs matching the physical inputs; absent optional inputs give None."""

    means: Tensor
    quats: Tensor
    scales: Tensor
    densities: Tensor
    beta: Optional[Tensor] = None
    uncertainty: Optional[Tensor] = None


class _VersionGuard:
    """The version counter of every distinct tensor a backward will read,
    taken when the forward returns; ``verify`` names the first one that an
    in-place operation has changed since."""

    def __init__(self, named):
        self._watched = []
        ids = set()
        for name, value in named:
            if not isinstance(value, Tensor) or id(value) in ids:
                continue
            ids.add(id(value))
            try:
                self._watched.append((name, value, value._version))
            except RuntimeError as error:  # inference tensors carry no version counter
                raise ValueError("explicit contexts require tensors with version counters") from error

    def verify(self) -> None:
        for name, value, version in self._watched:
            if value._version != version:
                raise RuntimeError(f"saved input {name} was modified in-place after forward")


class _Context:
    """What a forward hands its backward: the validated primitives, the saved
    buffers and options. Used once; the backward refuses a context of the
    other operator, a consumed one, or one whose inputs changed in place."""

    def __init__(self, kind: str, shape: tuple, primitives: tuple, buffers: dict, inputs, eps: float,
                 backward_mode: str, detach_variance: bool):
        self.kind = kind
        self.shape = shape
        self.primitives = primitives
        self.buffers = buffers
        self.eps = eps
        self.backward_mode = backward_mode
        self.detach_variance = detach_variance
        self.used = False
        self._guard = _VersionGuard(list(inputs) + list(zip(_PRIMITIVE_NAMES, primitives)) + list(buffers.items()))

    @staticmethod
    def open(context, kind: str) -> "_Context":
        if not isinstance(context, _Context) or context.kind != kind:
            raise TypeError(f"expected a context returned by {kind}_fwd")
        if context.used:
            raise RuntimeError("this context has already been consumed by backward")
        context._guard.verify()
        return context

    def adjoints(self, named) -> list:
        """Validates the output adjoints (None stays None) with one value check."""
        checks = check.ValueChecks()
        out = []
        for name, value in named:
            if value is not None:
                value = check.tensor(name, value, self.shape, self.primitives[0])
                checks.finite(name, value)
            out.append(value)
        checks.run()
        return out

    def zero_gradients(self) -> ParameterGradients:
        return ParameterGradients(*(None if p is None else torch.zeros_like(p) for p in self.primitives))

    def close(self, gradients: ParameterGradients) -> ParameterGradients:
        """Marks the context consumed and drops everything it holds."""
        self.used = True
        self.primitives = ()
        self.buffers.clear()
        self._guard = _VersionGuard(())
        return gradients


def _mode(value) -> str:
    value = str(value).lower()
    if value not in ("pixel", "primitive"):
        raise ValueError("backward_mode must be 'pixel' or 'primitive'")
    return value


def _cameras(viewmats, Ks, means, height: int, width: int, checks: check.ValueChecks):
    """viewmats [C,4,4] (world-to-camera, last row 0 0 0 1) and pinhole Ks
    [C,3,3] (positive focal lengths, zero skew, last row 0 0 1), constant."""
    for name, value in (("viewmats", viewmats), ("Ks", Ks)):
        check.constant(name, value)
    viewmats = check.tensor("viewmats", viewmats, (None, 4, 4), means)
    C = viewmats.shape[0]
    if C == 0:
        raise ValueError("viewmats must contain at least one camera")
    if C > 65535 or C * max(height * width, means.shape[0]) > check.INT32_MAX:
        raise ValueError("camera batch exceeds the kernels' indexing or grid limits")
    if max(height, width) > 16 * 65535:
        raise ValueError("image_size exceeds the kernels' grid limits")
    Ks = check.tensor("Ks", Ks, (C, 3, 3), means)
    checks.finite("viewmats", viewmats)
    checks.finite("Ks", Ks)
    checks.element("Ks must have positive fx and fy", Ks, be.CHECK_K_FOCAL, 9)
    checks.element("Ks must be pinhole intrinsics with zero skew and last row [0, 0, 1]", Ks, be.CHECK_K_PINHOLE, 9)
    checks.element("viewmats must have homogeneous last row [0, 0, 0, 1]", viewmats, be.CHECK_VIEW_ROW, 16)
    return viewmats, Ks


_FRAME_CACHE: dict = {}


def _constant_vector(value, means):
    """A volume-frame 3-vector given as Python numbers: (in the primitives'
    dtype, in float64 as given, nonfinite?, nonpositive?) with the flags of
    the rounded vector, built once per value, device and dtype (the tensors
    are never written), or None for any other form of the argument."""
    if not (isinstance(value, (tuple, list)) and len(value) == 3 and all(type(v) in (int, float) for v in value)):
        return None
    try:
        # hex keys tell -0.0 from 0.0 and match NaN
        key = (tuple(float(v).hex() for v in value), means.device, means.dtype)
    except OverflowError:
        return None
    hit = _FRAME_CACHE.get(key)
    if hit is None:
        try:
            rounded = torch.as_tensor(value, dtype=means.dtype)
            exact = torch.as_tensor(value, dtype=torch.float64)
        except (OverflowError, RuntimeError, TypeError):
            return None
        if len(_FRAME_CACHE) >= 64:
            _FRAME_CACHE.clear()
        hit = (rounded.to(means.device), exact.to(means.device), not bool(torch.isfinite(rounded).all()),
               bool((rounded <= 0).any()))
        _FRAME_CACHE[key] = hit
    return hit


def _volume_frame(extent, center, means, checks: check.ValueChecks):
    """volume_extent (positive) and volume_center: 3-vectors, constant. ->
    (extent, center) in the primitives' dtype, and both in float64 as given
    (Python numbers are not rounded to float32 there). Python-number vectors
    are checked on the host (in the same order as device values)."""
    out, exact, nonpositive = [], [], None
    for name, value in (("volume_extent", extent), ("volume_center", center)):
        check.constant(name, value)
        given = value
        hit = _constant_vector(value, means)
        if hit is not None:
            out.append(hit[0])
            exact.append(hit[1])
            checks.known(f"{name} must contain only finite values", hit[2])
            if name == "volume_extent":
                nonpositive = hit[3]
            continue
        if not isinstance(value, Tensor):
            value = torch.as_tensor(value, device=means.device, dtype=means.dtype)
        value = check.tensor(name, value, (3,), means)
        checks.finite(name, value)
        out.append(value)
        if isinstance(given, Tensor):
            exact.append(value.double().contiguous())
        else:
            exact.append(torch.as_tensor(given, device=means.device, dtype=torch.float64).reshape(3).contiguous())
    message = "volume_extent must contain positive physical lengths"
    if nonpositive is None:
        checks.element(message, out[0], be.CHECK_NONPOSITIVE)
    else:
        checks.known(message, nonpositive)
    return out[0], out[1], exact[0], exact[1]


# ---------------------------------------------------------------------------
# projection
# ---------------------------------------------------------------------------
_STATE_KEEP = ("viewmats", "Ks", "extent", "center", "extent64", "center64")


def _drop_state(buffers, eps, radius_clip):
    for key in [k for k in buffers if k not in _STATE_KEEP]:
        del buffers[key]
    buffers["recompute"] = (eps, radius_clip)


def _raster_state(prims, viewmats, Ks, width, height, eps, radius_clip) -> dict:
    """Projection and tile binning: everything the rasterizer and its VJP read."""
    means, quats, scales, densities, beta, uncertainty = prims
    C = viewmats.shape[0]
    if _use_fused(means):
        f = fused.prepare(means, quats, scales, densities, viewmats, Ks, width, height, eps, radius_clip, beta=beta,
                          uncertainty=uncertainty)
        # the waves per tile of the configuration for this shape (the tile
        # edge stays the operators' 16)
        _, f.fwd_waves, f.bwd_waves = fused.config_for_shape(means.shape[0], C, height, width, beta is not None)
        return {"fused": f}
    T = RASTER_TILE
    with stage("project"):
        covars, _ = ops.parameter_transform_fwd(quats, scales)
        mapped, cov2d, radii = ops.world_to_image_projection_fwd(means, covars, viewmats, Ks, width, height, eps,
                                                                 radius_clip)
        conics, base_mus = ops.conics_mu_fwd(cov2d, eps)
    opacity = densities.unsqueeze(0).expand(C, -1).contiguous()
    mus, beta2d, ratio = base_mus, None, None
    if beta is not None:
        ratio = beta_path_ratio(beta)
        mus = base_mus * ratio.unsqueeze(0)
        beta2d = (beta + 0.5).unsqueeze(0).expand(C, -1).contiguous()
    tiles_x, tiles_y = (width + T - 1) // T, (height + T - 1) // T
    with stage("bin"):
        _, isect_ids, flatten_ids = ops.isect_tiles(mapped, radii, T, tiles_x, tiles_y)
        offsets = ops.isect_offset_encode(isect_ids, C, tiles_x, tiles_y)
    return dict(covars=covars, mapped=mapped, cov2d=cov2d, conics=conics, base_mus=base_mus, mus=mus,
                opacity=opacity, beta2d=beta2d, ratio=ratio, offsets=offsets, isect_ids=isect_ids,
                flatten_ids=flatten_ids)


def _raster_render(b, prims, shape):
    means, _, _, _, beta, uncertainty = prims
    C, height, width = shape
    if "fused" in b:
        return fused.render(b["fused"], C, means.device, b["fused"].fwd_waves)
    T = RASTER_TILE
    prefix = "beta_" if beta is not None else ""
    args = (b["mapped"], b["conics"], b["opacity"], b["mus"]) + (() if beta is None else (b["beta2d"],))
    with stage("rasterize_fwd"):
        if uncertainty is None:
            return getattr(ops, prefix + "rasterize_to_pixels_fwd")(*args, b["offsets"], b["flatten_ids"], width,
                                                                    height, T)
        return getattr(ops, prefix + "rasterize_to_pixels_mean_var_fwd")(
            *args, uncertainty, b["offsets"], b["flatten_ids"], width, height, T)
@torch.no_grad()
def rasterize_fwd(
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
    recompute: bool = False,
):
    """Projects primitives to CT images [C, H, W] (or (mean, variance)) and a VJP context.

    ``recompute=True`` keeps only the inputs in the context and re-runs the
    projection and tile binning in the backward instead of holding the
    per-intersection lists between the two calls (for camera batches whose
    intersection lists do not fit next to the rest of the working set).
    """
    inputs = (("means", means), ("quats", quats), ("scales", scales), ("densities", densities),
              ("viewmats", viewmats), ("Ks", Ks), ("beta", beta), ("uncertainty", uncertainty))
    height, width = check.positive_ints("image_size", image_size, 2)
    values = check.ValueChecks()
    prims = check.primitives((means, quats, scales, densities, beta, uncertainty), values)
    means, quats, scales, densities, beta, uncertainty = prims
    eps, radius_clip = check.options(beta, eps, radius_clip, detach_variance)
    mode = _mode(backward_mode)
    viewmats, Ks = _cameras(viewmats, Ks, means, height, width, values)
    values.run()
    C = viewmats.shape[0]
    shape = (C, height, width)
    buffers = {"viewmats": viewmats, "Ks": Ks}
    if means.shape[0] == 0:
        image = means.new_zeros(shape)
        output = image if uncertainty is None else (image, image.clone())
    else:
        buffers.update(_raster_state(prims, viewmats, Ks, width, height, eps, radius_clip))
        output = _raster_render(buffers, prims, shape)
        if recompute:
            _drop_state(buffers, eps, radius_clip)
    return output, _Context("rasterize", shape, prims, buffers, inputs, eps, mode, detach_variance)


@torch.no_grad()
def rasterize_bwd(context: _Context, v_image: Optional[Tensor], v_variance: Optional[Tensor] = None):
    """Consumes a projection context; returns physical-parameter gradients."""
    context = _Context.open(context, "rasterize")
    means, quats, scales, densities, beta, uncertainty = context.primitives
    if v_variance is not None and uncertainty is None:
        raise ValueError("v_variance requires uncertainty in the forward call")
    v_image, v_variance = context.adjoints((("v_image", v_image), ("v_variance", v_variance)))
    if means.shape[0] == 0 or (v_image is None and v_variance is None):
        return context.close(context.zero_gradients())
    if v_image is None:
        v_image = means.new_zeros(context.shape)
    b = context.buffers
    if "recompute" in b:
        _, height, width = context.shape
        b.update(_raster_state(context.primitives, b["viewmats"], b["Ks"], width, height, *b.pop("recompute")))
    if "fused" in b:
        grads = fused.backward(b["fused"], means, quats, scales, densities, b["viewmats"], b["Ks"], v_image,
                               waves=b["fused"].bwd_waves, v_variance=v_variance,
                               detach_variance=context.detach_variance)
        v_means, v_quats, v_scales, v_densities = grads[:4]
        v_beta = grads[4] if beta is not None else None
        v_uncertainty = grads[-1] if uncertainty is not None else None
        return context.close(ParameterGradients(v_means, v_quats, v_scales, v_densities, v_beta, v_uncertainty))
    _, height, width = context.shape
    T = RASTER_TILE
    prefix = "beta_" if beta is not None else ""
    args = (b["mapped"], b["conics"], b["opacity"], b["mus"]) + (() if beta is None else (b["beta2d"],))
    with stage("rasterize_bwd"):
        if v_variance is not None:
            grads = getattr(ops, prefix + "rasterize_to_pixels_mean_var_bwd_per_gaussian")(
                *args, b["offsets"], b["isect_ids"], b["flatten_ids"], v_image, v_variance, uncertainty,
                context.detach_variance, width, height, T, False)
            v_uncertainty = grads[-1]
        else:
            if context.backward_mode == "primitive":
                grads = getattr(ops, prefix + "rasterize_to_pixels_bwd_per_gaussian")(
                    *args, b["offsets"], b["isect_ids"], b["flatten_ids"], v_image, width, height, T, False)
            else:
                grads = getattr(ops, prefix + "rasterize_to_pixels_bwd")(
                    *args, b["offsets"], b["flatten_ids"], v_image, width, height, T, False)
            v_uncertainty = None if uncertainty is None else torch.zeros_like(uncertainty)
    v_mapped, v_conics, v_opacity, v_mus = grads[1:5]
    v_beta = None
    if beta is not None:
        # mu_beta = mu_gauss * ratio(beta) and the projected exponent is beta + 1/2.
        v_beta = grads[5].sum(dim=0) + (v_mus * b["mus"]).sum(dim=0) * beta_path_log_grad(beta)
        v_mus = (v_mus * b["ratio"].unsqueeze(0)).contiguous()
    with stage("project_bwd"):
        v_cov2d = ops.conics_mu_bwd(b["cov2d"], context.eps, b["conics"], b["base_mus"], v_conics.contiguous(),
                                    v_mus.contiguous())
        v_means, v_covars = ops.world_to_image_projection_bwd(means, b["covars"], b["viewmats"], b["Ks"], width,
                                                              height, v_mapped.contiguous(), v_cov2d)
        v_quats, v_scales = ops.parameter_transform_bwd(quats, scales, v_covars)
        v_densities = v_opacity.sum(dim=0)
    return context.close(ParameterGradients(v_means, v_quats, v_scales, v_densities, v_beta, v_uncertainty))


# ---------------------------------------------------------------------------
# voxelization
# ---------------------------------------------------------------------------
def _voxel_state(prims, vol: dict, shape, eps, radius_clip) -> dict:
    """Volume mapping and 3D tile binning (vol: the volume frame, extent /
    center in the primitives' dtype and extent64 / center64 in float64)."""
    means, quats, scales, densities, beta, uncertainty = prims
    extent, center = vol["extent"], vol["center"]
    depth, height, width = shape
    T = VOXEL_TILE
    with stage("voxel_map"):
        covars, max_scales = ops.parameter_transform_fwd(quats, scales)
        mapped, conics, radii = ops.world_to_volume_mapping_fwd(means, covars, max_scales, extent, center, width,
                                                                height, depth, eps, radius_clip)
    tx, ty, tz = (width + T - 1) // T, (height + T - 1) // T, (depth + T - 1) // T
    with stage("voxel_bin"):
        _, isect_ids, flatten_ids, tile_cum = _ops_impl._isect_tiles_3d(mapped, radii, T, tx, ty, tz)
        offsets = ops.isect_offset_encode_3d(isect_ids, tx, ty, tz)
    state = dict(covars=covars, mapped=mapped, conics=conics, radii=radii, tile_cum=tile_cum, offsets=offsets,
                 isect_ids=isect_ids, flatten_ids=flatten_ids)
    if _use_fused_voxels(means):
        with stage("voxel_map"):
            state["frame"] = fused.voxel_frame(means, quats, scales, vol["extent64"], vol["center64"],
                                               (width, height, depth), eps)
    return state


def _voxel_render(b, prims, shape):
    means, quats, scales, densities, beta, uncertainty = prims
    depth, height, width = shape
    T = VOXEL_TILE
    prefix = "beta_" if beta is not None else ""
    args = (b["mapped"], b["conics"], densities) + (() if beta is None else (beta,))
    with stage("voxelize_fwd"):
        if _use_fused_voxels(means):
            return fused.voxel_render(b, densities, beta, (width, height, depth), T, uncertainty=uncertainty)
        if uncertainty is None:
            return getattr(ops, prefix + "voxelize_to_voxels_fwd")(*args, b["offsets"], b["flatten_ids"], width,
                                                                   height, depth, T)
        return getattr(ops, prefix + "voxelize_to_voxels_mean_var_fwd")(
            *args, uncertainty, b["offsets"], b["flatten_ids"], width, height, depth, T)


@torch.no_grad()
def voxelize_fwd(
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
    recompute: bool = False,
):
    """Samples primitives at voxel centres: volume [D, H, W] (or (mean, variance)) and a VJP context.

    ``recompute=True`` re-runs the volume mapping and 3D binning in the
    backward instead of holding the intersection lists in the context.
    """
    inputs = (("means", means), ("quats", quats), ("scales", scales), ("densities", densities),
              ("volume_extent", volume_extent), ("volume_center", volume_center), ("beta", beta),
              ("uncertainty", uncertainty))
    depth, height, width = check.positive_ints("volume_shape", volume_shape, 3)
    if max(height, width) > 8 * 65535:
        raise ValueError("volume_shape exceeds the kernels' grid limits")
    values = check.ValueChecks()
    prims = check.primitives((means, quats, scales, densities, beta, uncertainty), values)
    means, quats, scales, densities, beta, uncertainty = prims
    eps, radius_clip = check.options(beta, eps, radius_clip, detach_variance)
    mode = _mode(backward_mode)
    extent, center, extent64, center64 = _volume_frame(volume_extent, volume_center, means, values)
    values.run()
    shape = (depth, height, width)
    buffers = {"extent": extent, "center": center, "extent64": extent64, "center64": center64}
    if means.shape[0] == 0:
        volume = means.new_zeros(shape)
        output = volume if uncertainty is None else (volume, volume.clone())
    else:
        buffers.update(_voxel_state(prims, buffers, shape, eps, radius_clip))
        output = _voxel_render(buffers, prims, shape)
        if recompute:
            _drop_state(buffers, eps, radius_clip)
    return output, _Context("voxelize", shape, prims, buffers, inputs, eps, mode, detach_variance)


@torch.no_grad()
def voxelize_bwd(context: _Context, v_volume: Optional[Tensor], v_variance: Optional[Tensor] = None):
    """Consumes a volume context; returns physical-parameter gradients."""
    context = _Context.open(context, "voxelize")
    means, quats, scales, densities, beta, uncertainty = context.primitives
    if v_variance is not None and uncertainty is None:
        raise ValueError("v_variance requires uncertainty in the forward call")
    v_volume, v_variance = context.adjoints((("v_volume", v_volume), ("v_variance", v_variance)))
    if means.shape[0] == 0 or (v_volume is None and v_variance is None):
        return context.close(context.zero_gradients())
    if v_volume is None:
        v_volume = means.new_zeros(context.shape)
    b = context.buffers
    if "recompute" in b:
        b.update(_voxel_state(context.primitives, b, context.shape, *b.pop("recompute")))
    depth, height, width = context.shape
    T = VOXEL_TILE
    prefix = "beta_" if beta is not None else ""
    args = (b["mapped"], b["conics"], densities) + (() if beta is None else (beta,))
    with stage("voxelize_bwd"):
        if _use_fused_voxels(means):
            grads = fused.voxel_vjp(b, densities, beta, v_volume, (width, height, depth), T, uncertainty=uncertainty,
                                    v_variance=v_variance, detach_variance=context.detach_variance)
            if v_variance is not None:
                v_uncertainty = grads[-1]
            else:
                v_uncertainty = None if uncertainty is None else torch.zeros_like(uncertainty)
        elif v_variance is not None:
            grads = getattr(ops, prefix + "voxelize_to_voxels_mean_var_bwd_per_gaussian")(
                *args, b["offsets"], b["isect_ids"], b["flatten_ids"], v_volume, v_variance, uncertainty,
                context.detach_variance, width, height, depth, T)
            v_uncertainty = grads[-1]
        else:
            if context.backward_mode == "primitive":
                grads = getattr(ops, prefix + "voxelize_to_voxels_bwd_per_gaussian")(
                    *args, b["offsets"], b["isect_ids"], b["flatten_ids"], v_volume, width, height, depth, T)
            else:
                grads = getattr(ops, prefix + "voxelize_to_voxels_bwd")(
                    *args, b["offsets"], b["flatten_ids"], v_volume, width, height, depth, T)
            v_uncertainty = None if uncertainty is None else torch.zeros_like(uncertainty)
    v_mapped, v_conics, v_densities = grads[:3]
    v_beta = None if beta is None else grads[3]
    with stage("voxel_map_bwd"):
        v_means, v_covars = ops.world_to_volume_mapping_bwd(means, b["covars"], b["extent"], b["center"], width,
                                                            height, depth, context.eps, v_mapped.contiguous(),
                                                            v_conics.contiguous())
        v_quats, v_scales = ops.parameter_transform_bwd(quats, scales, v_covars)
    return context.close(ParameterGradients(v_means, v_quats, v_scales, v_densities, v_beta, v_uncertainty))


__all__ = ["ParameterGradients", "rasterize_fwd", "rasterize_bwd", "voxelize_fwd", "voxelize_bwd"]

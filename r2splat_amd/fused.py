"""Fused float32 Gaussian / Beta rasterization (8, 16 or 32 pixel tiles) on the HIP library.

``FusedForward`` holds what the backward needs (the tile lists' entries and
the projected records, tile offsets, per-(camera, position) intersection
offsets and the primitive order, device counters) for one forward call. The backward stores one partial per intersection and
sums them per primitive in a fixed order, so it is bitwise reproducible. Capacity for the
intersection list is sized from the previous call with the same shapes
(``CapacityCache``) and grown on overflow, so steady-state calls perform no
host-visible sizing besides the overflow check.

Beta primitives (``beta`` given) carry the per-primitive line-integral factor
ratio(beta) and d log ratio / d beta, computed once per forward in float64
as r2splat/_beta.py does and cast to float32 (explicit.py:166-170, 226-230).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import torch

from ._beta import beta_path_log_grad, beta_path_ratio
from .hip import _backend as be
from .hip._backend import call, ptr
from .profiling import stage

TILE = 16  # default tile edge (the explicit API's)


class CapacityCache:
    """Per (device, C, N, H, W) intersection capacity, grown on overflow."""

    def __init__(self, headroom: float = 1.15):
        self.headroom = headroom
        self._cap = {}

    def get(self, key, C, N):
        return self._cap.get(key, max(1024, 4 * C * N))

    def known(self, key) -> bool:
        return key in self._cap

    def grow(self, key, needed: int):
        self._cap[key] = max(1024, int(math.ceil(needed * self.headroom)))
        return self._cap[key]

    def settle(self, key, used: int):
        cap = self._cap.get(key)
        if cap is None or used > cap or cap > 2 * used * self.headroom + 4096:
            self.grow(key, used)


CAPACITY = CapacityCache()

# Device counters of the most recent prepare() ({intersections, overflow});
# lets a caller of captured steps check the overflow flag after replaying.
LAST_COUNTERS = None


@dataclass
class FusedForward:
    packed: torch.Tensor
    offsets: torch.Tensor
    tile_order: torch.Tensor
    record_offsets: torch.Tensor
    counters: torch.Tensor
    capacity: int
    n_isects: int
    width: int
    height: int
    eps: float
    tile_size: int = TILE
    beta: Optional[torch.Tensor] = None       # [N] (Beta primitives) or None (Gaussian)
    ratio: Optional[torch.Tensor] = None      # [N] B(1/2, beta + 1) / sqrt(2 pi)
    log_grad: Optional[torch.Tensor] = None   # [N] d log ratio / d beta
    uncertainty: Optional[torch.Tensor] = None  # [N]: render also returns the variance
    # waves per tile of the render and of the raster VJP that callers choose
    # for this forward (0: the kernels' default, 8)
    fwd_waves: int = 0
    bwd_waves: int = 0

    @property
    def is_beta(self) -> int:
        return int(self.beta is not None)

    @property
    def kind(self) -> int:
        """Record kind of the C ABI: bit 0 Beta, bit 1 uncertainty."""
        return _kind(self.beta, self.uncertainty)


def config_for_shape(n: int, views: int, height: int, width: int, beta: bool = False) -> tuple:
    """(tile edge, forward waves, backward waves) from a job's shape alone:
    the configuration the autotuner (``batched.Autotuner``) picks most often
    for jobs of that shape. Measured on MI355X over 56 jobs of the serving mix
    (N 10k-400k, 16-128 views of 128^2-512^2, Gaussian and Beta), its prepare
    + render + raster VJP time is within 1% of tuning every job and 5.5%
    below tile 16 with 8 waves. Gaussians on 512^2 images take 32-pixel tiles
    once views x N reaches 500k; Beta primitives on 512^2 images measured
    fastest with 16-pixel tiles; small images take 16 waves per tile."""
    side = min(height, width)
    if side >= 512:
        if beta:
            return (16, 4, 4)
        return (32 if views * n >= 500_000 else 16, 8 if views <= 32 and n >= 60_000 else 4, 4)
    if side >= 256:
        return (16, 8 if views <= 16 else 4, 4 if beta else 8)
    return (16, 16, 16) if views <= 32 else (16, 8, 16)


def _kind(beta, uncertainty) -> int:
    return int(beta is not None) | (int(uncertainty is not None) << 1)


def _scratch(C, N, width, height, tile_size, capacity, beta, device):
    nbytes = be.lib().r2s_fused_scratch_bytes(C, N, width, height, tile_size, capacity, beta)
    return torch.empty((nbytes,), dtype=torch.uint8, device=device)


def beta_factors(beta: torch.Tensor):
    """(ratio, log_grad) of float32 betas, as the explicit path computes them."""
    return beta_path_ratio(beta).contiguous(), beta_path_log_grad(beta).contiguous()


def prepare(means, quats, scales, densities, viewmats, Ks, width, height, eps, radius_clip, *, capacity=None,
            check=None, tile_size: int = TILE, beta=None, factors=None, uncertainty=None) -> FusedForward:
    """Primitive order, projection and binning into tile lists (everything but the image). float32, one device.

    ``check`` (default: unless the stream is being captured into a graph)
    reads the overflow flag and re-runs with more capacity. Under capture the
    capacity of the last eager call with the same shapes is used and the flag
    is left on the device (``FusedForward.counters[1]``) for the caller.
    ``beta`` [N] selects Beta primitives; ``factors`` = beta_factors(beta)
    may be passed in to reuse them; ``uncertainty`` [N] adds the variance."""
    device = means.device
    C, N = viewmats.shape[0], means.shape[0]
    kind = _kind(beta, uncertainty)
    ratio = log_grad = None
    if beta is not None:
        ratio, log_grad = factors if factors is not None else beta_factors(beta)
    key = (device, C, N, height, width, tile_size, kind)
    capturing = torch.cuda.is_current_stream_capturing()
    if check is None:
        check = not capturing
    if capturing and capacity is None and not CAPACITY.known(key):
        raise RuntimeError("fused.prepare under graph capture needs a capacity: run one eager call with the "
                           "same shapes first")
    cap = int(capacity or CAPACITY.get(key, C, N))
    ntx, nty = -(-width // tile_size), -(-height // tile_size)
    offsets = torch.empty((C, nty, ntx), dtype=torch.int32, device=device)
    tile_order = torch.empty((C * nty * ntx,), dtype=torch.int32, device=device)
    # per-record intersection offsets (C*N + 1) followed by the primitive order (N)
    record_offsets = torch.empty((C * N + 1 + N,), dtype=torch.int32, device=device)
    counters = torch.zeros((2,), dtype=torch.int32, device=device)
    with torch.cuda.device(device):
        stream = be.stream_of(device)
        n_isects = -1
        while True:
            scratch = _scratch(C, N, width, height, tile_size, cap, kind, device)
            packed = torch.empty((be.lib().r2s_fused_packed_bytes(C, N, cap, kind),), dtype=torch.uint8,
                                 device=device)
            with stage("project_bin"):
                call("r2s_fused_prepare", ptr(means), ptr(quats), ptr(scales), ptr(densities), ptr(beta), ptr(ratio),
                     ptr(uncertainty), ptr(viewmats), ptr(Ks), C, N, width, height, tile_size, float(eps), float(radius_clip), cap, ptr(scratch),
                     scratch.numel(), ptr(packed), ptr(offsets), ptr(tile_order), ptr(record_offsets), ptr(counters),
                     stream)
            del scratch
            if not check:
                break
            n_isects, overflow = (int(v) for v in counters.tolist())
            if not overflow:
                CAPACITY.settle(key, n_isects)
                break
            cap = CAPACITY.grow(key, n_isects)
    global LAST_COUNTERS
    LAST_COUNTERS = counters
    return FusedForward(packed, offsets, tile_order, record_offsets, counters, cap, n_isects, width, height,
                        float(eps), tile_size, beta, ratio, log_grad, uncertainty)


def render(fwd: FusedForward, C: int, device, waves: int = 0):
    """The image [C, H, W] of a prepared forward, or (image, variance) when it
    carries uncertainties (`waves` per tile: 4, 8, 16; 0 = 8)."""
    image = torch.empty((C, fwd.height, fwd.width), dtype=torch.float32, device=device)
    var = torch.empty_like(image) if fwd.uncertainty is not None else None
    with torch.cuda.device(device):
        with stage("rasterize_fwd"):
            call("r2s_fused_render", ptr(fwd.packed), ptr(fwd.offsets), ptr(fwd.tile_order), ptr(fwd.counters),
                 fwd.capacity, C, fwd.width, fwd.height, fwd.tile_size, waves, fwd.kind, ptr(image), ptr(var),
                 be.stream_of(device))
    return image if var is None else (image, var)


def forward(means, quats, scales, densities, viewmats, Ks, width, height, eps, radius_clip, *, capacity=None,
            check=None, tile_size: int = TILE, waves: int = 0, beta=None, uncertainty=None):
    """Renders [C, H, W] (with uncertainty: (image, variance)); returns
    (output, FusedForward). float32, contiguous, one device."""
    fwd = prepare(means, quats, scales, densities, viewmats, Ks, width, height, eps, radius_clip,
                  capacity=capacity, check=check, tile_size=tile_size, beta=beta, uncertainty=uncertainty)
    return render(fwd, viewmats.shape[0], means.device, waves), fwd


def backward(fwd: FusedForward, means, quats, scales, densities, viewmats, Ks, v_image, waves: int = 0,
             v_variance=None, detach_variance: bool = False):
    """VJP of the fused forward -> (v_means, v_quats, v_scales, v_densities[, v_beta]
    [, v_uncertainty]); v_beta for Beta primitives, v_uncertainty (and the
    variance adjoint ``v_variance``, whose path to the mean parameters
    ``detach_variance`` cuts) for forwards with uncertainties."""
    device = means.device
    C, N = viewmats.shape[0], means.shape[0]
    grad = torch.empty((fwd.capacity * 8,), dtype=torch.float32, device=device)
    ws = torch.empty((be.lib().r2s_fused_param_workspace_bytes(C, N),), dtype=torch.uint8, device=device)
    v_means = torch.empty_like(means)
    v_quats = torch.empty_like(quats)
    v_scales = torch.empty_like(scales)
    v_dens = torch.empty_like(densities)
    v_beta = torch.empty_like(fwd.beta) if fwd.is_beta else None
    v_unc = torch.empty_like(fwd.uncertainty) if fwd.uncertainty is not None else None
    with torch.cuda.device(device):
        stream = be.stream_of(device)
        with stage("rasterize_bwd"):
            call("r2s_fused_raster_vjp", ptr(fwd.packed), ptr(fwd.offsets), ptr(fwd.tile_order), ptr(fwd.counters),
                 fwd.capacity, C,
                 fwd.width, fwd.height, fwd.tile_size, waves, fwd.kind, ptr(v_image), ptr(v_variance),
                 int(bool(detach_variance)), ptr(grad), stream)
        with stage("param_bwd"):
            call("r2s_fused_param_vjp", ptr(means), ptr(quats), ptr(scales), ptr(densities), ptr(fwd.beta),
                 ptr(fwd.ratio), ptr(fwd.log_grad), ptr(viewmats), ptr(Ks), C, N, fwd.width, fwd.height, fwd.eps,
                 ptr(fwd.record_offsets), ptr(fwd.counters), fwd.capacity, ptr(grad), ptr(ws), ws.numel(),
                 ptr(v_means), ptr(v_quats), ptr(v_scales), ptr(v_dens), ptr(v_beta), ptr(v_unc), stream)
    out = (v_means, v_quats, v_scales, v_dens)
    if fwd.is_beta:
        out = out + (v_beta,)
    if v_unc is not None:
        out = out + (v_unc,)
    return out


def voxel_frame(means, quats, scales, extent, center, dims, eps: float) -> torch.Tensor:
    """Per-primitive voxel-frame mean and inverse covariance in double
    precision ([N, 9] float64: mean x y z, conic xx xy xz yy yz zz), from
    float32 parameters and the volume frame (extent [3], center [3], float64)."""
    N = means.shape[0]
    W, H, D = dims
    frame = torch.empty((N, 9), dtype=torch.float64, device=means.device)
    with torch.cuda.device(means.device):
        call("r2s_voxel_frame", ptr(means), ptr(quats), ptr(scales), ptr(extent), ptr(center), W, H, D, float(eps),
             N, ptr(frame), be.stream_of(means.device))
    return frame


def voxel_vjp(state: dict, densities, beta, v_volume, dims, tile_size: int = 8, *, uncertainty=None,
              v_variance=None, detach_variance: bool = False):
    """Per-primitive voxelize VJP (float32, 8^3 tiles) from a voxel binning
    state (mapped, conics, radii, tile_cum, isect_ids, flatten_ids, and the
    double-precision frame of voxel_frame): one partial
    per intersection, summed per primitive in a fixed order (no atomics).
    With a variance adjoint (``v_variance``, and the forward's ``uncertainty``)
    it also returns v_uncertainty.
    -> (v_mapped [N,3], v_conics [N,6], v_densities [N][, v_beta [N]][, v_uncertainty [N]])."""
    mapped = state["mapped"]
    device = mapped.device
    N = mapped.shape[0]
    W, H, D = dims
    T = int(tile_size)
    ids, flat = state["isect_ids"], state["flatten_ids"]
    n_isects = ids.numel()
    partials = torch.empty((max(1, be.lib().r2s_voxel_partials_bytes(n_isects)),), dtype=torch.uint8, device=device)
    v_mapped = torch.empty((N, 3), dtype=torch.float32, device=device)
    v_conics = torch.empty((N, 6), dtype=torch.float32, device=device)
    v_dens = torch.empty((N,), dtype=torch.float32, device=device)
    v_beta = torch.empty((N,), dtype=torch.float32, device=device) if beta is not None else None
    v_unc = torch.empty((N,), dtype=torch.float32, device=device) if v_variance is not None else None
    with torch.cuda.device(device):
        call("r2s_voxelize_fused_vjp", ptr(mapped), ptr(state["frame"]), ptr(state["conics"]),
             ptr(densities), ptr(beta),
             ptr(uncertainty if v_variance is not None else None), ptr(state["radii"]), ptr(state["tile_cum"]),
             ptr(ids), ptr(flat), n_isects, N, W, H, D, T, -(-W // T), -(-H // T), -(-D // T), ptr(v_volume),
             ptr(v_variance), int(bool(detach_variance)), ptr(partials), partials.numel(), ptr(v_mapped),
             ptr(v_conics), ptr(v_dens), ptr(v_beta), ptr(v_unc), be.stream_of(device))
    out = (v_mapped, v_conics, v_dens)
    if beta is not None:
        out = out + (v_beta,)
    if v_unc is not None:
        out = out + (v_unc,)
    return out


def voxel_render(state: dict, densities, beta, dims, tile_size: int = 8, *, uncertainty=None):
    """Volume [D, H, W] of a voxel binning state (float32, 8^3 tiles, with the
    double-precision frame of voxel_frame), or (volume, variance) with
    ``uncertainty`` [N]."""
    mapped = state["mapped"]
    device = mapped.device
    W, H, D = dims
    T = int(tile_size)
    out = torch.empty((D, H, W), dtype=torch.float32, device=device)
    var = torch.empty_like(out) if uncertainty is not None else None
    with torch.cuda.device(device):
        call("r2s_voxelize_fused_fwd", ptr(mapped), ptr(state["frame"]), ptr(state["conics"]),
             ptr(densities), ptr(beta),
             ptr(uncertainty), ptr(state["offsets"]), ptr(state["flatten_ids"]), state["flatten_ids"].numel(),
             mapped.shape[0], W, H, D, T, -(-W // T), -(-H // T), -(-D // T), ptr(out), ptr(var),
             be.stream_of(device))
    return out if var is None else (out, var)

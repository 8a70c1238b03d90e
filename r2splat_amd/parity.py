"""Float64 evaluation of the fused float32 path's own (list entry, pixel) pairs.

A precision audit of the fused float32 rasterizer. It keeps a pair when
its float32 alpha passes the 1e-5 cut-off; a float64 evaluation of the same
scene can decide a pair at the threshold the other way (each such term is
about 1e-5). To hold the float32 arithmetic to the contract's bounds on its
own, this module evaluates in float64, over exactly the pairs of the fused
path's tile lists:

* ``kept``: the pairs the float32 kernels kept (their keep masks from
  r2s_fused_keep_masks: the forward's for images, the backward's for
  gradients), with float64 projections (the staged operator path in float64:
  means2d, conics, mu, opacity);
* ``ref``: the float64 path's own decisions (sigma >= 0 and |alpha| >= 1e-5,
  Beta 0 <= d^2 < 1). Over the float64 path's own tile lists (``fp64_table``)
  this is the float64 reference itself; the two paths' lists can differ where
  a footprint radius rounds across an integer (``coverage_diff``).

Per-pair VJPs follow the reference's formulas (RasterizeToPixelsBwd.cu:
Gaussian sigma = 1/2 (a dx^2 + c dy^2) + b dx dy with dx = pixel - mean; Beta
d^2 = delta^T P delta with delta = mean - pixel and alpha = w (1 - d^2)^b2d;
variance sum q alpha^2, whose adjoint reaches alpha unless detached) and are
chained through the float64 conic/mu, projection and parameter-transform VJP
operators exactly as the explicit API chains them. ``ref`` reproduces the
float64 operator path's image and gradients (checked by the GPU tests), which
pins this evaluation itself. Used by the float64 parity tests and by the
benchmark's parity report.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch

from . import explicit
from . import hip as ops
from ._beta import beta_path_log_grad
from .hip import _backend as be
from .hip._backend import ptr

MODES = ("kept", "ref")


@dataclass
class PairTable:
    """Per list entry of the fused path: camera, input primitive index and
    tile coordinates (int64, on the device), plus the keep masks."""

    cam: torch.Tensor
    prim: torch.Tensor
    ty: torch.Tensor
    tx: torch.Tensor
    mask_fwd: Optional[torch.Tensor]  # int32 [n, T*T/32] bit words (None: a table with no float32 decisions)
    mask_bwd: Optional[torch.Tensor]
    tile: int


def pair_table(fwd, C: int, N: int) -> PairTable:
    """Entries of a fused forward (r2splat_amd.fused.FusedForward) and its keep masks."""
    T = fwd.tile_size
    n = int(fwd.counters[0].item())
    assert int(fwd.counters[1].item()) == 0, "the forward overflowed its capacity"
    words = T * T // 32
    dev = fwd.packed.device
    mask_f = torch.empty((max(n, 1), words), dtype=torch.int32, device=dev)
    mask_b = torch.empty_like(mask_f)
    be.call("r2s_fused_keep_masks", ptr(fwd.packed), ptr(fwd.offsets), ptr(fwd.counters), fwd.capacity, C,
            fwd.width, fwd.height, T, fwd.kind, ptr(mask_f), ptr(mask_b), be.stream_of(dev))
    entries = fwd.packed[: 8 * n].view(torch.int32).view(n, 2)
    rec = entries[:, 0].long()
    perm = fwd.record_offsets[C * N + 1: C * N + 1 + N].long()
    cam = rec // N
    prim = perm[rec % N]
    offsets = fwd.offsets.reshape(-1).long()
    tile = torch.searchsorted(offsets, torch.arange(n, device=dev), right=True) - 1
    ntx = -(-fwd.width // T)
    nty = -(-fwd.height // T)
    t = tile % (ntx * nty)
    assert bool((tile // (ntx * nty) == cam).all()), "an entry sits in another camera's list"
    return PairTable(cam, prim, t // ntx, t % ntx, mask_f[:n], mask_b[:n], T)


def list_table(buffers: dict, C: int, N: int, width: int, height: int, tile: int) -> PairTable:
    """Entries of the staged operator path's tile lists (isect_ids = camera *
    n_tiles + ty * n_tiles_x + tx, flatten_ids = camera * N + primitive), with
    no keep masks: a table for mode 'ref' on a path's own coverage."""
    ids = buffers["isect_ids"].long()
    flat = buffers["flatten_ids"].long()
    ntx = -(-width // tile)
    nty = -(-height // tile)
    cam = ids // (ntx * nty)
    t = ids % (ntx * nty)
    assert bool((cam == flat // N).all()), "an entry's key and primitive name different cameras"
    return PairTable(cam, flat % N, t // ntx, t % ntx, None, None, tile)


def coverage_diff(a: PairTable, b: PairTable, C: int, N: int) -> int:
    """(camera, primitive) records whose tile counts differ between two tables."""
    ca = torch.bincount(a.cam * N + a.prim, minlength=C * N)
    cb = torch.bincount(b.cam * N + b.prim, minlength=C * N)
    return int((ca != cb).sum())


def _bits(words: torch.Tensor, T: int) -> torch.Tensor:
    shifts = torch.arange(32, device=words.device, dtype=torch.int32)
    return ((words[:, :, None] >> shifts) & 1).reshape(words.shape[0], T * T).bool()


@dataclass
class Fp64Scene:
    """The float64 operator path's state for the same inputs (explicit API buffers)."""

    ctx: object
    buffers: dict  # the forward's buffers (the context's own are dropped when its backward consumes it)
    image: object
    means: torch.Tensor
    quats: torch.Tensor
    scales: torch.Tensor
    beta: Optional[torch.Tensor]
    uncertainty: Optional[torch.Tensor]
    viewmats: torch.Tensor
    Ks: torch.Tensor
    shape: tuple
    eps: float


def fp64_scene(means, quats, scales, densities, viewmats, Ks, image_size, *, beta=None, uncertainty=None,
               eps=1e-8, detach=False) -> Fp64Scene:
    d = lambda t: None if t is None else t.double().contiguous()  # noqa: E731
    args = [d(means), d(quats), d(scales), d(densities), d(viewmats), d(Ks)]
    out, ctx = explicit.rasterize_fwd(*args, image_size, beta=d(beta), uncertainty=d(uncertainty), eps=eps,
                                      detach_variance=detach)
    assert "fused" not in ctx.buffers
    return Fp64Scene(ctx, dict(ctx.buffers), out, args[0], args[1], args[2], d(beta), d(uncertainty), args[4], args[5],
                     (viewmats.shape[0],) + tuple(image_size), eps)


def fp64_table(sc: "Fp64Scene") -> PairTable:
    """The float64 operator path's own tile lists as a table (mode 'ref')."""
    C, H, W = sc.shape
    return list_table(sc.buffers, C, sc.means.shape[0], W, H, explicit.RASTER_TILE)


@dataclass
class PairResult:
    image: torch.Tensor
    variance: Optional[torch.Tensor]
    grads: Optional[dict]
    flips_fwd: int = 0  # pairs whose float32 forward decision differs from the float64 one
    flips_bwd: int = 0
    pairs: int = 0      # in-image pairs evaluated


def evaluate(table: PairTable, sc: Fp64Scene, mode: str, v_image=None, v_variance=None, detach=False,
             chunk: int = 1 << 19) -> PairResult:
    """Image (and variance) and, when adjoints are given, the float64
    gradients (means, quats, scales, densities[, beta][, uncertainty]) over the
    table's pairs, kept by the float32 masks (mode 'kept') or by the float64
    decisions (mode 'ref')."""
    assert mode in MODES
    assert mode == "ref" or table.mask_fwd is not None, "mode 'kept' needs a table with keep masks"
    b = sc.buffers
    C, H, W = sc.shape
    T = table.tile
    N = sc.means.shape[0]
    dev = sc.means.device
    is_beta = sc.beta is not None
    var_on = sc.uncertainty is not None
    m2 = b["mapped"].reshape(C * N, 2)
    con = b["conics"].reshape(C * N, 3)
    opac = b["opacity"].reshape(C * N)
    mus = b["mus"].reshape(C * N)
    b2d = b["beta2d"].reshape(C * N) if is_beta else None
    unc = sc.uncertainty
    image = torch.zeros(C * H * W, dtype=torch.float64, device=dev)
    variance = torch.zeros_like(image) if var_on else None
    backward = v_image is not None or v_variance is not None
    if backward:
        vi = (v_image if v_image is not None else torch.zeros(sc.shape, device=dev)).double().reshape(-1)
        vv = v_variance.double().reshape(-1) if v_variance is not None else None
        acc = torch.zeros((C * N, 8), dtype=torch.float64, device=dev)  # mx my a b c opac mu b2d
        acc_q = torch.zeros(N, dtype=torch.float64, device=dev)
    r = torch.arange(T * T, device=dev) // T
    col = torch.arange(T * T, device=dev) % T
    flips_f = flips_b = pairs = 0
    n = table.cam.numel()
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        rec = table.cam[lo:hi] * N + table.prim[lo:hi]
        x = table.tx[lo:hi, None] * T + col
        y = table.ty[lo:hi, None] * T + r
        inimg = (x < W) & (y < H)
        pix = (table.cam[lo:hi, None] * H + y.clamp_max(H - 1)) * W + x.clamp_max(W - 1)
        mx, my = m2[rec, 0:1], m2[rec, 1:2]
        a, bb, c = con[rec, 0:1], con[rec, 1:2], con[rec, 2:3]
        w = (opac[rec] * mus[rec])[:, None]
        if is_beta:
            dx, dy = mx - (x + 0.5), my - (y + 0.5)
            d2 = a * dx * dx + 2 * bb * dx * dy + c * dy * dy
            inside = (d2 >= 0) & (d2 < 1)
            h = torch.where(inside, 1 - d2, torch.ones_like(d2))
            bexp = b2d[rec][:, None]
            e = torch.where(inside, h.pow(bexp), torch.zeros_like(d2))  # alpha / w
            alpha = w * e
            keep64 = inside & (alpha.abs() >= 1e-5)
        else:
            dx, dy = (x + 0.5) - mx, (y + 0.5) - my
            sigma = 0.5 * (a * dx * dx + c * dy * dy) + bb * dx * dy
            e = torch.exp(-sigma)
            alpha = w * e
            keep64 = (sigma >= 0) & (alpha.abs() >= 1e-5)
        keep64 &= inimg
        pairs += int(inimg.sum())
        keep_f = keep_b = keep64
        if table.mask_fwd is not None:
            kf = _bits(table.mask_fwd[lo:hi], T) & inimg
            kb = _bits(table.mask_bwd[lo:hi], T) & inimg
            flips_f += int((kf != keep64).sum())
            flips_b += int((kb != keep64).sum())
            if mode == "kept":
                keep_f, keep_b = kf, kb
        af = torch.where(keep_f, alpha, torch.zeros_like(alpha))
        image.index_add_(0, pix.reshape(-1), af.reshape(-1))
        if var_on:
            variance.index_add_(0, pix.reshape(-1), (unc[table.prim[lo:hi]][:, None] * af * af).reshape(-1))
        if not backward:
            continue
        G = vi[pix]
        ab = torch.where(keep_b, alpha, torch.zeros_like(alpha))
        eb = torch.where(keep_b, e, torch.zeros_like(e))
        if var_on and vv is not None:
            vvp = vv[pix]
            q = unc[table.prim[lo:hi]][:, None]
            if not detach:
                G = G + 2.0 * q * ab * vvp
            acc_q.index_add_(0, table.prim[lo:hi], (vvp * ab * ab).sum(1))
        G = torch.where(keep_b, G, torch.zeros_like(G))
        v_w = (G * eb).sum(1)
        if is_beta:
            hb = torch.where(keep_b, h, torch.ones_like(h))
            v_d2 = -G * w * bexp * eb / hb  # d alpha / d d2 = -w b2d h^(b2d - 1)
            parts = [(v_d2 * 2 * (a * dx + bb * dy)).sum(1), (v_d2 * 2 * (bb * dx + c * dy)).sum(1),
                     (v_d2 * dx * dx).sum(1), (v_d2 * 2 * dx * dy).sum(1), (v_d2 * dy * dy).sum(1)]
            v_b2d = (G * ab * torch.log(hb)).sum(1)
        else:
            v_sig = -G * ab
            parts = [(-v_sig * (a * dx + bb * dy)).sum(1), (-v_sig * (bb * dx + c * dy)).sum(1),
                     (0.5 * v_sig * dx * dx).sum(1), (v_sig * dx * dy).sum(1), (0.5 * v_sig * dy * dy).sum(1)]
            v_b2d = torch.zeros_like(v_w)
        cols = torch.stack(parts + [v_w * mus[rec], v_w * opac[rec], v_b2d], 1)
        acc.index_add_(0, rec, cols)
    img = image.reshape(sc.shape)
    var = variance.reshape(sc.shape) if var_on else None
    grads = _chain(sc, acc, acc_q, v_variance is not None) if backward else None
    return PairResult(img, var, grads, flips_f, flips_b, pairs)


def _chain(sc: Fp64Scene, acc: torch.Tensor, acc_q: torch.Tensor, with_var: bool) -> dict:
    """Per-(camera, primitive) VJPs -> parameter gradients, as explicit.rasterize_bwd chains them."""
    b = sc.buffers
    C, H, W = sc.shape
    N = sc.means.shape[0]
    acc = acc.reshape(C, N, 8)
    v_mapped = acc[..., 0:2].contiguous()
    v_conics = acc[..., 2:5].contiguous()
    v_opacity = acc[..., 5]
    v_mus = acc[..., 6]
    out = {}
    if sc.beta is not None:
        out["beta"] = acc[..., 7].sum(0) + (v_mus * b["mus"]).sum(0) * beta_path_log_grad(sc.beta)
        v_mus = v_mus * b["ratio"][None]
    v_cov2d = ops.conics_mu_bwd(b["cov2d"], sc.eps, b["conics"], b["base_mus"], v_conics, v_mus.contiguous())
    v_means, v_covars = ops.world_to_image_projection_bwd(sc.means, b["covars"], b["viewmats"], b["Ks"], W, H,
                                                          v_mapped, v_cov2d)
    v_quats, v_scales = ops.parameter_transform_bwd(sc.quats, sc.scales, v_covars)
    out.update(means=v_means, quats=v_quats, scales=v_scales, densities=v_opacity.sum(0))
    if sc.uncertainty is not None and with_var:
        out["uncertainty"] = acc_q
    return out

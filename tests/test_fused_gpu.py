"""Fused float32 pipeline: parity with the oracle and with the staged operator
path, capacity overflow handling, concurrent contexts, edge tiles, and
working-size agreement (100k Gaussians x 64 views of 256^2)."""

import numpy as np
import pytest
import torch

import r2splat_amd
from oracle import splat as O
from r2splat_amd import explicit, fused, workloads
from r2splat_amd import hip as ops
from r2splat_amd.hip import _backend as be

from r2splat_amd.parity import evaluate, fp64_scene, pair_table

pytestmark = pytest.mark.gpu


def _run(sc, backend, v_fn=None):
    old = explicit.BACKEND
    explicit.BACKEND = backend
    try:
        img, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size)
        v = v_fn(img) if v_fn else (2.0 * (img - 0.8 * img) / img.numel()).contiguous()
        return img, r2splat_amd.rasterize_bwd(ctx, v), v, ctx
    finally:
        explicit.BACKEND = old


def _proj64(sc, clip=3.0, W=None, H=None, eps=1e-8):
    """means2d and radii of the operators in float64: the tile coverage the
    fused projection (which projects in double) reproduces."""
    covars, _ = ops.parameter_transform_fwd(sc.quats.double(), sc.scales.double())
    m2, _, radii = ops.world_to_image_projection_fwd(sc.means.double(), covars, sc.viewmats.double(), sc.Ks.double(),
                                                     W or sc.width, H or sc.height, eps, clip)
    return m2, radii


def _mask(sc, eps=1e-8, clip=3.0):
    m2, radii = _proj64(sc, clip, eps=eps)
    return O.tile_mask_2d(m2.cpu().numpy(), radii.cpu().numpy(), 16, sc.width, sc.height)


@pytest.mark.parametrize("size", [(64, 64), (50, 37)])
def test_fused_matches_oracle(device, size):
    W, H = size
    sc = workloads.scene(600, 3, H, W, scale_range=(0.02, 0.1), seed=4, device=device)
    img, grads, v, ctx = _run(sc, "fused")
    assert "fused" not in ctx.buffers  # consumed
    leaves = [t.detach().cpu().double().requires_grad_() for t in sc.params()]
    ref = O.rasterize_scene(*leaves, sc.viewmats.cpu().double(), sc.Ks.cpu().double(), (H, W), eps=1e-8,
                            mask=_mask(sc))
    assert O.relative_error(img, ref) <= 1e-5
    exp = torch.autograd.grad((ref * v.cpu().double()).sum(), leaves)
    for name, e in zip(("means", "quats", "scales", "densities"), exp):
        err = O.relative_error(getattr(grads, name), e)
        assert err <= 1e-4, f"{name}: {err:.2e}"


def test_fused_matches_staged_ops_path(device):
    """Config 0: the fused pipeline (projection in double, records and
    rasterization in float32) and the staged float32 operators agree to 1e-5
    on the image and 1e-4 on every gradient."""
    sc = workloads.config0(device=device)
    a_img, a, _, _ = _run(sc, "fused")
    b_img, b, _, _ = _run(sc, "ops")
    assert O.relative_error(a_img, b_img) <= 1e-5
    for name in ("means", "quats", "scales", "densities"):
        assert O.relative_error(getattr(a, name), getattr(b, name)) <= 1e-4, name


def _count_grows(monkeypatch):
    calls = []
    grow = fused.CAPACITY.grow
    monkeypatch.setattr(fused.CAPACITY, "grow", lambda key, needed: calls.append(needed) or grow(key, needed))
    return calls


def test_fused_forward_is_deterministic_and_grows_capacity(device, monkeypatch):
    sc = workloads.config0(device=device)
    key = (sc.means.device, 4, 1000, 64, 64, 16, 0)  # (device, C, N, H, W, tile, record kind)
    fused.CAPACITY._cap[key] = 1024  # force an overflow on the next call
    grows = _count_grows(monkeypatch)
    img1, f1 = fused.forward(*sc.params(), sc.viewmats, sc.Ks, 64, 64, 1e-8, 3.0)
    assert grows == [f1.n_isects]  # the first pass overflowed and the capacity was regrown
    assert f1.capacity >= f1.n_isects > 1024
    assert int(f1.counters[1].item()) == 0
    img2, f2 = fused.forward(*sc.params(), sc.viewmats, sc.Ks, 64, 64, 1e-8, 3.0)
    assert torch.equal(img1, img2)
    assert f2.n_isects == f1.n_isects
    counts, ids, flat = ops.isect_tiles(*_proj(sc), 16, 4, 4)
    assert ids.numel() == f1.n_isects


def test_sort_path_overflow_is_flagged_and_regrown(device, monkeypatch):
    """The sort + pack binning (500 tiles of 8 pixels per camera, more bins
    than the direct-placement table holds) under a forced capacity overflow:
    the forward flags it, regrows, and renders bitwise the same image as a
    call that never overflowed."""
    sc = workloads.scene(2000, 2, 160, 200, scale_range=(0.01, 0.06), seed=12, device=device)
    args = (*sc.params(), sc.viewmats, sc.Ks, 200, 160, 1e-8, 3.0)
    ref, fr = fused.forward(*args, tile_size=8, capacity=1 << 22)
    assert int(fr.counters[1].item()) == 0 and fr.n_isects > 256
    key = (sc.means.device, 2, 2000, 160, 200, 8, 0)  # (device, C, N, H, W, tile, record kind)
    fused.CAPACITY._cap[key] = 256  # force an overflow on the next call
    grows = _count_grows(monkeypatch)
    img, f = fused.forward(*args, tile_size=8)
    assert grows == [fr.n_isects]  # the first pass overflowed and the capacity was regrown
    assert f.capacity >= f.n_isects == fr.n_isects > 256
    assert int(f.counters[1].item()) == 0
    assert torch.equal(img, ref)
    # a capacity given explicitly and too small raises the device flag and renders nothing
    prep = fused.prepare(*args, tile_size=8, capacity=256, check=False)
    assert int(prep.counters[1].item()) == 1
    assert not bool(fused.render(prep, 2, sc.means.device).any())


def _proj(sc):
    return _proj64(sc, 3.0)


def test_fused_tile_offsets_match_operator_binning(device):
    sc = workloads.scene(3000, 5, 96, 80, scale_range=(0.01, 0.08), seed=9, device=device)
    img, f = fused.forward(*sc.params(), sc.viewmats, sc.Ks, 80, 96, 1e-8, 3.0)
    counts, ids, _ = ops.isect_tiles(*_proj(sc), 16, 5, 6)
    off = ops.isect_offset_encode(ids, 5, 5, 6)
    assert f.n_isects == ids.numel()
    assert torch.equal(f.offsets, off)
    # the primitive order (after the offsets): a permutation, sorted by the Morton code of the means
    CN = counts.numel()
    perm = f.record_offsets[CN + 1:].long()
    assert torch.equal(torch.sort(perm).values, torch.arange(3000, device=device))
    lo, hi = sc.means.min(0).values, sc.means.max(0).values
    q = ((sc.means - lo) / (hi - lo) * 32).clamp(0, 31).long()
    code = torch.zeros(3000, dtype=torch.long, device=device)
    for bit in range(5):
        for k in range(3):
            code |= ((q[:, k] >> bit) & 1) << (3 * bit + 2 - k)
    code = code >> 7  # sets of up to 2^17 primitives sort by the code's top 8 bits
    assert bool((code[perm][1:] >= code[perm][:-1]).all())
    # ties keep the input order (the sort is stable)
    tie = code[perm][1:] == code[perm][:-1]
    assert bool((perm[1:][tie] > perm[:-1][tie]).all())
    # per-(camera, position) intersection offsets: exclusive scan of the tile-rectangle areas in that order
    flat = counts[:, perm].reshape(-1).long()
    expect = torch.zeros(flat.numel() + 1, dtype=torch.long, device=device)
    expect[1:] = torch.cumsum(flat, 0)
    assert torch.equal(f.record_offsets[:CN + 1].long(), expect)


@pytest.mark.parametrize("W,H,T", [(80, 96, 16), (520, 504, 16), (40, 24, 8)])
def test_tile_order_is_banded_longest_first(device, W, H, T):
    """The tile order (direct placement; 33 x 32 tiles per camera: sort and
    pack): a permutation of the tiles in which XCD band x (per + (x < rem)
    tiles of the camera-major index from x per + min(x, rem)) holds exactly
    its own tiles, in non-increasing size class floor(8 log2(count + 1)):
    each tile's count + 1 is below 2^(1/8) times its predecessor's."""
    sc = workloads.scene(2500, 3, H, W, scale_range=(0.01, 0.08), seed=13, device=device)
    f = fused.prepare(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, 3.0, tile_size=T)
    n = f.tile_order.numel()
    order = f.tile_order.long()
    assert torch.equal(torch.sort(order).values, torch.arange(n, device=device))
    off = f.offsets.reshape(-1).long()
    ends = torch.cat([off[1:], torch.tensor([f.n_isects], device=device)])
    size = (ends - off).double() + 1
    per, rem = n // 8, n % 8
    for x in range(8):
        lo = x * per + min(x, rem)
        hi = lo + per + (1 if x < rem else 0)
        band = order[lo:hi]
        assert bool(((band >= lo) & (band < hi)).all()), x
        assert bool((size[band][1:] <= size[band][:-1] * 2 ** 0.125 * (1 + 1e-6)).all()), x


@pytest.mark.parametrize("use_beta", [False, True])
def test_packed_lists_hold_each_emission_once_in_its_tile(device, use_beta):
    """The packed tile lists are a permutation of the emission order: each
    emission index (g field) appears exactly once, in the camera-major list of
    its own (camera, tile), on a record of its own primitive (same projected
    mean as the staged projection)."""
    sc = workloads.scene(3000, 5, 96, 80, scale_range=(0.01, 0.08), seed=9, device=device,
                         beta_range=(0.5, 4.0) if use_beta else None)
    clip = 1.0 if use_beta else 3.0
    _, f = fused.forward(*sc.params(), sc.viewmats, sc.Ks, 80, 96, 1e-8, clip, beta=sc.beta)
    n, C, N, ntx, nty = f.n_isects, 5, 3000, 5, 6
    # stored record width in floats (Beta records carry their double-precision form after the 12 floats)
    width = be.lib().r2s_fused_record_bytes(f.kind) // 4
    assert width == (24 if use_beta else 8)
    # packed = list entries {record index, emission index}, then the C*N records
    entries = f.packed[: n * 8].view(torch.int32).reshape(n, 2).cpu().numpy()
    rec0 = -(-f.capacity * 8 // 256) * 256
    records = f.packed[rec0: rec0 + C * N * width * 4].view(torch.float32).reshape(C * N, width).cpu().numpy()
    g = entries[:, 1]
    mean = records[entries[:, 0], :2]
    assert np.array_equal(np.sort(g), np.arange(n))
    # records follow the primitive order stored after the offsets
    perm = f.record_offsets[C * N + 1:].long().cpu().numpy()
    m2, radii = _proj_clip(sc, clip)
    m2n = m2.cpu().numpy()[:, perm]
    x0, x1, y0, y1 = O.tile_rects_2d(m2n, radii.cpu().numpy()[:, perm], 16, ntx, nty)
    counts = ((x1 - x0) * (y1 - y0)).reshape(-1)
    cam = np.repeat(np.arange(C), N)
    keys, owners = O._emit(counts, [y0.reshape(-1), x0.reshape(-1)], [(y1 - y0).reshape(-1), (x1 - x0).reshape(-1)],
                           [ntx, 1], cam * (ntx * nty))
    offsets = f.offsets.reshape(-1).cpu().numpy().astype(np.int64)
    ends = np.append(offsets[1:], n)
    lists = np.repeat(np.arange(offsets.size), ends - offsets)
    assert np.array_equal(keys[g], lists)  # every record sits in the list of its own tile
    assert np.array_equal(entries[:, 0], owners[g])  # and names the record that emitted it
    np.testing.assert_allclose(mean, m2n.reshape(-1, 2)[owners[g]], rtol=1e-5, atol=1e-3)


def test_outstanding_contexts_are_independent(device):
    a = workloads.scene(500, 2, 48, 48, scale_range=(0.02, 0.1), seed=1, device=device)
    b = workloads.scene(700, 2, 48, 48, scale_range=(0.02, 0.1), seed=2, device=device)
    ia, ca = r2splat_amd.rasterize_fwd(*a.params(), a.viewmats, a.Ks, a.image_size)
    ib, cb = r2splat_amd.rasterize_fwd(*b.params(), b.viewmats, b.Ks, b.image_size)
    gb = r2splat_amd.rasterize_bwd(cb, torch.ones_like(ib))
    ga = r2splat_amd.rasterize_bwd(ca, torch.ones_like(ia))
    ia2, ca2 = r2splat_amd.rasterize_fwd(*a.params(), a.viewmats, a.Ks, a.image_size)
    ga2 = r2splat_amd.rasterize_bwd(ca2, torch.ones_like(ia2))
    assert torch.equal(ia, ia2)
    # the fused backward sums per-intersection partials in a fixed order: bitwise reproducible
    for name in ("means", "quats", "scales", "densities"):
        assert torch.equal(getattr(ga, name), getattr(ga2, name)), name


def test_working_size_fused_agrees_with_float64_operator_path(device):
    """Config 1 (100k Gaussians, 64 views of 256^2): the fused float32 pipeline
    against the staged operator path in float64 (itself pinned to the oracle
    at config 0), everything included (pairs decided differently at the 1e-5
    cut-off too): 1e-5 on the image and 1e-4 on every gradient. The fused
    projection computes radii and tile rectangles in double, so the two paths
    cover the same tiles."""
    sc = workloads.config1(device=device)
    tgt = workloads.config1(device=device, seed=1)
    target, _ = r2splat_amd.rasterize_fwd(*tgt.params(), tgt.viewmats, tgt.Ks, tgt.image_size)
    vfn = lambda img: ((2.0 / img.numel()) * (img - target)).contiguous()
    a_img, a, v, _ = _run(sc, "fused", vfn)
    p64 = [t.double() for t in sc.params()]
    img64, ctx64 = r2splat_amd.rasterize_fwd(*p64, sc.viewmats.double(), sc.Ks.double(), sc.image_size)
    b = r2splat_amd.rasterize_bwd(ctx64, v.double())
    err = O.relative_error(a_img, img64)
    print(f"config1 fused float32 vs float64 operators: image {err:.2e}", flush=True)
    assert err <= 1e-5
    for name in ("means", "quats", "scales", "densities"):
        err = O.relative_error(getattr(a, name), getattr(b, name))
        print(f"  {name}: {err:.2e}", flush=True)
        assert err <= 1e-4, f"{name}: {err:.2e}"
    # additivity over a split of the primitive set (each primitive's cut-off depends only on
    # itself, so render(A u B) = render(A) + render(B) up to summation order)
    half = sc.means.shape[0] // 2
    parts = []
    for sl in (slice(0, half), slice(half, None)):
        p = [t[sl].contiguous() for t in sc.params()]
        parts.append(r2splat_amd.rasterize_fwd(*p, sc.viewmats, sc.Ks, sc.image_size)[0])
    assert O.relative_error(parts[0] + parts[1], a_img) <= 1e-5
    # the backward is linear in the adjoint image: doubling it doubles every gradient bitwise
    img, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size)
    g2 = r2splat_amd.rasterize_bwd(ctx, (2.0 * vfn(img)).contiguous())
    for name in ("means", "quats", "scales", "densities"):
        assert torch.equal(getattr(g2, name), 2.0 * getattr(a, name)), name


# 72 x 56: every tile size bins by direct placement; 200 x 160 at 8-pixel tiles
# (500 tiles per camera, more bins than the placement table holds) sorts and packs
@pytest.mark.parametrize("tile,waves,W,H", [(8, 4, 72, 56), (8, 16, 72, 56), (32, 4, 72, 56), (32, 16, 72, 56),
                                           (16, 16, 72, 56), (8, 8, 200, 160)])
def test_fused_tile_sizes_and_wave_counts_match_oracle(device, tile, waves, W, H):
    sc = workloads.scene(700, 3, H, W, scale_range=(0.02, 0.1), seed=6, device=device)
    img, f = fused.forward(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, 3.0, tile_size=tile, waves=waves)
    _, ids, _ = ops.isect_tiles(*_proj(sc), tile, -(-W // tile), -(-H // tile))
    assert f.n_isects == ids.numel()
    v = (2.0 * (img - 0.8 * img) / img.numel()).contiguous()
    grads = fused.backward(f, *sc.params(), sc.viewmats, sc.Ks, v, waves=waves)
    m2, radii = _proj64(sc, 3.0, W, H)
    mask = O.tile_mask_2d(m2.cpu().numpy(), radii.cpu().numpy(), tile, W, H)
    leaves = [t.detach().cpu().double().requires_grad_() for t in sc.params()]
    ref = O.rasterize_scene(*leaves, sc.viewmats.cpu().double(), sc.Ks.cpu().double(), (H, W), eps=1e-8, mask=mask)
    assert O.relative_error(img, ref) <= 1e-5
    exp = torch.autograd.grad((ref * v.cpu().double()).sum(), leaves)
    for name, g, e in zip(("means", "quats", "scales", "densities"), grads, exp):
        err = O.relative_error(g, e)
        assert err <= 1e-4, f"{name}: {err:.2e}"
    # the backward does not depend on the wave count; the forward only through summation order
    img8, f8 = fused.forward(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, 3.0, tile_size=tile, waves=8)
    assert O.relative_error(img8, img) <= 1e-6
    grads8 = fused.backward(f8, *sc.params(), sc.viewmats, sc.Ks, v, waves=8)
    for a, b in zip(grads, grads8):
        assert torch.equal(a, b)


# --------------------------------------------------------------------------- Beta primitives
def _beta_scene(device, n=700, c=3, H=56, W=72, seed=6):
    return workloads.scene(n, c, H, W, scale_range=(0.02, 0.1), seed=seed, device=device, beta_range=(0.5, 4.0))


def _beta_oracle(sc, v, tile, clip=1.0):
    m2, radii = _proj64(sc, clip)
    mask = O.tile_mask_2d(m2.cpu().numpy(), radii.cpu().numpy(), tile, sc.width, sc.height)
    leaves = [t.detach().cpu().double().requires_grad_() for t in sc.params() + (sc.beta,)]
    ref = O.rasterize_scene(*leaves[:4], sc.viewmats.cpu().double(), sc.Ks.cpu().double(), sc.image_size, eps=1e-8,
                            beta=leaves[4], mask=mask)
    return ref, torch.autograd.grad((ref * v.cpu().double()).sum(), leaves)


@pytest.mark.parametrize("tile,waves,W,H", [(16, 8, 72, 56), (8, 4, 72, 56), (32, 16, 72, 56), (8, 8, 200, 160)])
def test_fused_beta_matches_oracle(device, tile, waves, W, H):
    """Beta primitives (alpha = w (1 - d^2)^(beta + 1/2), w = density mu ratio(beta)):
    image to 1e-5 and all five gradients, beta included, to 1e-4 of the fp64 oracle
    (both binning paths, as in the Gaussian test)."""
    sc = _beta_scene(device, H=H, W=W)
    W, H = sc.width, sc.height
    img, f = fused.forward(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, 1.0, tile_size=tile, waves=waves,
                           beta=sc.beta)
    _, ids, _ = ops.isect_tiles(*_proj_clip(sc, 1.0), tile, -(-W // tile), -(-H // tile))
    assert f.n_isects == ids.numel()
    v = (2.0 * (img - 0.8 * img) / img.numel()).contiguous()
    grads = fused.backward(f, *sc.params(), sc.viewmats, sc.Ks, v, waves=waves)
    assert len(grads) == 5
    ref, exp = _beta_oracle(sc, v, tile)
    assert O.relative_error(img, ref) <= 1e-5
    for name, g, e in zip(("means", "quats", "scales", "densities", "beta"), grads, exp):
        err = O.relative_error(g, e)
        assert err <= 1e-4, f"{name}: {err:.2e}"
    # the backward does not depend on the wave count
    img8, f8 = fused.forward(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, 1.0, tile_size=tile, waves=8,
                             beta=sc.beta)
    grads8 = fused.backward(f8, *sc.params(), sc.viewmats, sc.Ks, v, waves=8)
    for a, b in zip(grads, grads8):
        assert torch.equal(a, b)


def _proj_clip(sc, clip):
    return _proj64(sc, clip)


def _run_beta(sc, backend, v_fn=None):
    old = explicit.BACKEND
    explicit.BACKEND = backend
    try:
        img, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size, beta=sc.beta)
        v = v_fn(img) if v_fn else (2.0 * (img - 0.8 * img) / img.numel()).contiguous()
        return img, r2splat_amd.rasterize_bwd(ctx, v), v, ctx
    finally:
        explicit.BACKEND = old


@pytest.mark.parametrize("n,views,side,seed", [(1000, 4, 64, 0), (20_000, 8, 128, 3)])
def test_fused_beta_matches_float64_and_staged_ops_path(device, n, views, side, seed):
    """The explicit API's Beta path through the fused pipeline against float64
    over the pairs its kernels kept (r2splat_amd.parity): image within 1e-5,
    every gradient within 1e-4; bitwise reproducible backward. The staged
    float32 operators pass float32 means2d and conics between them (as the
    reference's operators do); the Beta profile's slope at its support edge
    (w b2d) carries that rounding into the image and gradients (measured up to
    1.5e-5 and 9e-5 at these sizes), so they are held to 3e-5 of the fused
    image and 2e-4 of its gradients."""
    sc = workloads.scene(n, views, side, side, scale_range=(0.005, 0.08), seed=seed, device=device,
                         beta_range=(0.5, 4.0))
    names = ("means", "quats", "scales", "densities", "beta")
    C, N = sc.viewmats.shape[0], sc.means.shape[0]
    img, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size, beta=sc.beta)
    assert "fused" in ctx.buffers
    table = pair_table(ctx.buffers["fused"], C, N)
    v = (2.0 * (img - 0.8 * img) / img.numel()).contiguous()
    a = r2splat_amd.rasterize_bwd(ctx, v)
    sc64 = fp64_scene(*sc.params(), sc.viewmats, sc.Ks, sc.image_size, beta=sc.beta)
    kept = evaluate(table, sc64, "kept", v_image=v)
    err = O.relative_error(img, kept.image)
    print(f"Beta {n} x {views} x {side}^2: image vs float64 on kept pairs {err:.2e}", flush=True)
    assert err <= 1e-5
    for name in names:
        err = O.relative_error(getattr(a, name), kept.grads[name])
        print(f"  {name}: {err:.2e}", flush=True)
        assert err <= 1e-4, f"{name}: {err:.2e}"
    b_img, b, _, _ = _run_beta(sc, "ops", lambda _: v)
    err = O.relative_error(b_img, img)
    print(f"  staged float32 operators vs fused: image {err:.2e}", flush=True)
    assert err <= 3e-5
    for name in names:
        err = O.relative_error(getattr(b, name), getattr(a, name))
        print(f"  staged {name}: {err:.2e}", flush=True)
        assert err <= 2e-4, f"staged {name}: {err:.2e}"
    img2, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size, beta=sc.beta)
    a2 = r2splat_amd.rasterize_bwd(ctx, v)
    assert torch.equal(img2, img)
    for name in names:
        assert torch.equal(getattr(a2, name), getattr(a, name)), name


def test_fused_beta_large_exponent_centre_pixel(device):
    """One Beta primitive with a huge exponent seen at its centre: alpha = w
    exactly (h = 1), d alpha / d beta from the line-integral factor only."""
    gamma, ratio, dlog = 1e5, 0.002236059592288547, -4.999962500249999e-6
    means = torch.tensor([[0., 0., 1.]], device=device)
    quats = means.new_tensor([[1., 0., 0., 0.]])
    scales = means.new_full((1, 3), 0.2)
    view = torch.eye(4, device=device)[None]
    K = means.new_tensor([[[1., 0., .5], [0., 1., .5], [0., 0., 1.]]])
    beta = means.new_tensor([gamma])
    img, f = fused.forward(means, quats, scales, means.new_tensor([1.]), view, K, 1, 1, 0.0, 1.0, beta=beta)
    g = fused.backward(f, means, quats, scales, means.new_tensor([1.]), view, K, torch.ones_like(img))
    expected = 0.2 * (2 * np.pi) ** 0.5 * ratio
    torch.testing.assert_close(img.cpu(), torch.full((1, 1, 1), expected), atol=0, rtol=4e-6)
    torch.testing.assert_close(g[3].cpu(), torch.tensor([expected]), atol=0, rtol=4e-6)
    torch.testing.assert_close(g[4].cpu(), torch.tensor([expected * dlog]), atol=0, rtol=4e-6)


# --------------------------------------------------------------------------- voxels
@pytest.mark.parametrize("use_beta,uncertainty,detach", [(False, False, False), (True, False, False),
                                                         (False, True, False), (True, True, False),
                                                         (True, True, True)])
def test_fused_voxel_vjp_matches_float64_and_is_reproducible(device, use_beta, uncertainty, detach):
    """float32 fused voxelize forward and backward (no atomics: one partial per
    intersection, summed per primitive in a fixed order), mean or (mean,
    variance) with detach_variance either way, on a volume whose edges cut
    tiles: volume (and variance) within 1e-5 and every gradient within 1e-4 of
    the float64 operator path; bitwise reproducible. The staged float32
    operators (which pass float32 conics between them and form d^2 in float32,
    as the reference's operators do) agree with the fused kernels to 1e-5 /
    1e-4 for Gaussians. For Beta primitives with beta < 1 the slope of
    (1 - d^2)^beta at the support edge (of h^(beta - 1) in the VJP) turns that
    rounding into 2e-5 on the volume and up to ~2e-2 on the gradients: their
    volume is held to 1e-4 of float64 and their gradient errors are printed."""
    sc = workloads.scene(20000, 1, 8, 8, scale_range=(0.005, 0.05), seed=2, device=device,
                         beta_range=(0.5, 4.0) if use_beta else None, uncertainty_max=0.1 if uncertainty else None)
    shape = (90, 77, 64)
    g = torch.Generator(device=device).manual_seed(3)
    v = torch.randn(shape, generator=g, device=device)
    vv = torch.randn(shape, generator=g, device=device) if uncertainty else None

    def run(backend, dtype=torch.float32):
        cast = lambda t: None if t is None else t.to(dtype)  # noqa: E731
        old = explicit.BACKEND
        explicit.BACKEND = backend
        try:
            out, ctx = r2splat_amd.voxelize_fwd(*(cast(t) for t in sc.params()), shape,
                                                volume_extent=(1.0, 1.1, 0.9), beta=cast(sc.beta),
                                                uncertainty=cast(sc.uncertainty), detach_variance=detach)
            return out, r2splat_amd.voxelize_bwd(ctx, cast(v), cast(vv))
        finally:
            explicit.BACKEND = old

    def outputs(out):
        return out if uncertainty else (out,)

    names = ("means", "quats", "scales", "densities") + (("beta",) if use_beta else ()) + \
        (("uncertainty",) if uncertainty else ())
    out_a, a = run("fused")
    out_64, g64 = run("fused", torch.float64)
    for x, y in zip(outputs(out_a), outputs(out_64)):
        assert O.relative_error(x, y) <= 1e-5
    for name in names:
        err = O.relative_error(getattr(a, name), getattr(g64, name))
        assert err <= 1e-4, f"{name}: {err:.2e}"
    out_b, b = run("ops")
    ref_out, ref_g = (outputs(out_64), g64) if use_beta else (outputs(out_a), a)
    for x, y in zip(outputs(out_b), ref_out):
        assert O.relative_error(x, y) <= (1e-4 if use_beta else 1e-5)
    for name in names:
        err = O.relative_error(getattr(b, name), getattr(ref_g, name))
        print(f"staged float32 operators, {name}: {err:.2e} from {'float64' if use_beta else 'fused'}", flush=True)
        assert use_beta or err <= 1e-4, f"staged {name}: {err:.2e}"
    out_a2, a2 = run("fused")
    for x, y in zip(outputs(out_a), outputs(out_a2)):
        assert torch.equal(x, y)
    for name in names:
        assert torch.equal(getattr(a, name), getattr(a2, name)), name


# --------------------------------------------------------------------------- uncertainty
def _var_scene(device, use_beta, n=700, c=3, H=56, W=72, seed=8):
    return workloads.scene(n, c, H, W, scale_range=(0.02, 0.1), seed=seed, device=device,
                           beta_range=(0.5, 4.0) if use_beta else None, uncertainty_max=0.5)


def _adjoint_pair(img, seed):
    g = torch.Generator(device=img.device).manual_seed(seed)
    return (torch.randn(img.shape, generator=g, device=img.device) / img.numel(),
            torch.randn(img.shape, generator=g, device=img.device) / img.numel())


@pytest.mark.parametrize("use_beta,detach,tile,waves", [(False, False, 16, 8), (True, False, 8, 4),
                                                        (False, True, 32, 16), (True, True, 16, 16),
                                                        (True, False, 32, 4)])
def test_fused_variance_matches_oracle(device, use_beta, detach, tile, waves):
    """(mean, variance = sum q alpha^2) to 1e-5 and every gradient, the
    uncertainty's included, to 1e-4 of the fp64 oracle under both adjoints;
    with detach_variance the variance adjoint reaches only the uncertainty."""
    sc = _var_scene(device, use_beta)
    W, H = sc.width, sc.height
    clip = 1.0 if use_beta else 3.0
    (img, var), f = fused.forward(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, clip, tile_size=tile, waves=waves,
                                  beta=sc.beta, uncertainty=sc.uncertainty)
    assert f.n_isects == ops.isect_tiles(*_proj_clip(sc, clip), tile, -(-W // tile), -(-H // tile))[1].numel()
    v, vv = _adjoint_pair(img, 11)
    grads = fused.backward(f, *sc.params(), sc.viewmats, sc.Ks, v, waves=waves, v_variance=vv,
                           detach_variance=detach)
    names = ("means", "quats", "scales", "densities") + (("beta",) if use_beta else ()) + ("uncertainty",)
    assert len(grads) == len(names)
    m2, radii = _proj_clip(sc, clip)
    mask = O.tile_mask_2d(m2.cpu().numpy(), radii.cpu().numpy(), tile, W, H)
    leaves = [t.detach().cpu().double().requires_grad_()
              for t in sc.params() + ((sc.beta,) if use_beta else ()) + (sc.uncertainty,)]
    ref, ref_var = O.rasterize_scene(*leaves[:4], sc.viewmats.cpu().double(), sc.Ks.cpu().double(), (H, W),
                                     eps=1e-8, beta=leaves[4] if use_beta else None, uncertainty=leaves[-1],
                                     detach=detach, mask=mask)
    assert O.relative_error(img, ref) <= 1e-5
    assert O.relative_error(var, ref_var) <= 1e-5
    loss = (ref * v.cpu().double()).sum() + (ref_var * vv.cpu().double()).sum()
    for name, g, e in zip(names, grads, torch.autograd.grad(loss, leaves)):
        err = O.relative_error(g, e)
        assert err <= 1e-4, f"{name}: {err:.2e}"
    # the mean and its VJP with the variance adjoint at zero are those of the mean-only kernels
    img_m, f_m = fused.forward(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, clip, tile_size=tile, waves=waves,
                               beta=sc.beta)
    assert O.relative_error(img, img_m) <= 1e-6
    g_m = fused.backward(f_m, *sc.params(), sc.viewmats, sc.Ks, v, waves=waves)
    (_, _), f0 = fused.forward(*sc.params(), sc.viewmats, sc.Ks, W, H, 1e-8, clip, tile_size=tile, waves=waves,
                               beta=sc.beta, uncertainty=sc.uncertainty)
    g0 = fused.backward(f0, *sc.params(), sc.viewmats, sc.Ks, v, waves=waves)
    for name, a, b in zip(names, g0, g_m):
        assert O.relative_error(a, b) <= 1e-6, name
    assert not bool(g0[-1].any())  # no variance adjoint: zero uncertainty gradient


@pytest.mark.parametrize("use_beta,detach", [(False, False), (True, False), (True, True)])
def test_fused_variance_matches_float64_on_kept_pairs(device, use_beta, detach):
    """The explicit API's (mean, variance) path through the fused pipeline (20k
    primitives, 8 views of 128^2) against float64 over the pairs the float32
    kernels kept (r2splat_amd.parity; float64 projections from the staged
    operators in float64, each pinned to the oracle separately): mean and
    variance within 1e-5, every gradient including the uncertainty's within
    1e-4 (max |a - b| / max |b|). The fused backward is bitwise reproducible
    and linear in the adjoints."""
    sc = workloads.scene(20_000, 8, 128, 128, scale_range=(0.005, 0.08), seed=3, device=device,
                         beta_range=(0.5, 4.0) if use_beta else None, uncertainty_max=0.1)
    names = ("means", "quats", "scales", "densities") + (("beta",) if use_beta else ()) + ("uncertainty",)
    C, N = sc.viewmats.shape[0], sc.means.shape[0]
    p32 = dict(zip(("means", "quats", "scales", "densities"), sc.params()), beta=sc.beta,
               uncertainty=sc.uncertainty, viewmats=sc.viewmats, Ks=sc.Ks)

    def run(p, scale=1.0, keep_table=False):
        (img, var), ctx = r2splat_amd.rasterize_fwd(p["means"], p["quats"], p["scales"], p["densities"],
                                                    p["viewmats"], p["Ks"], sc.image_size, beta=p["beta"],
                                                    uncertainty=p["uncertainty"], detach_variance=detach)
        assert "fused" in ctx.buffers
        table = pair_table(ctx.buffers["fused"], C, N) if keep_table else None
        v, vv = _adjoint_pair(img, 13)
        return img, var, r2splat_amd.rasterize_bwd(ctx, (scale * v).contiguous(), (scale * vv).contiguous()), \
            table, (v, vv)

    a_img, a_var, a, table, (v, vv) = run(p32, keep_table=True)
    sc64 = fp64_scene(*sc.params(), sc.viewmats, sc.Ks, sc.image_size, beta=sc.beta, uncertainty=sc.uncertainty,
                      detach=detach)
    kept = evaluate(table, sc64, "kept", v_image=v, v_variance=vv, detach=detach)
    assert O.relative_error(a_img, kept.image) <= 1e-5
    assert O.relative_error(a_var, kept.variance) <= 1e-5
    for name in names:
        err = O.relative_error(getattr(a, name), kept.grads[name])
        assert err <= 1e-4, f"{name}: {err:.2e}"
    c_img, c_var, c, _, _ = run(p32)
    assert torch.equal(c_img, a_img) and torch.equal(c_var, a_var)
    _, _, d, _, _ = run(p32, 2.0)
    for name in names:
        assert torch.equal(getattr(c, name), getattr(a, name)), name
        assert torch.equal(getattr(d, name), 2.0 * getattr(a, name)), name


@pytest.mark.parametrize("case", ["no_primitives", "all_off_screen", "single_tile_grid"])
def test_fused_float32_degenerate_scenes(device, case):
    """The float32 fused path through the explicit API on scenes with nothing
    to draw (no primitives; primitives whose footprints miss every image) and
    on an image of one tile: images and gradients match the float64 operator
    path (zero where nothing is drawn)."""
    H, W = (40, 48) if case != "single_tile_grid" else (12, 10)
    sc = workloads.scene(64, 3, H, W, scale_range=(0.01, 0.08), seed=21, device=device)
    means, quats, scales, dens = sc.params()
    if case == "no_primitives":
        means, quats, scales, dens = means[:0], quats[:0], scales[:0], dens[:0]
    elif case == "all_off_screen":
        means = means + torch.tensor([0.0, 50.0, 0.0], device=device)  # far above every orbit camera's view
    img, ctx = r2splat_amd.rasterize_fwd(means, quats, scales, dens, sc.viewmats, sc.Ks, (H, W))
    assert img.shape == (3, H, W) and img.dtype == torch.float32
    v = torch.randn((3, H, W), device=device, generator=torch.Generator(device=device).manual_seed(2))
    grads = r2splat_amd.rasterize_bwd(ctx, v)
    ref_img, ref_ctx = r2splat_amd.rasterize_fwd(means.double(), quats.double(), scales.double(), dens.double(),
                                                 sc.viewmats.double(), sc.Ks.double(), (H, W))
    ref = r2splat_amd.rasterize_bwd(ref_ctx, v.double())
    if case != "single_tile_grid":
        assert torch.count_nonzero(img) == 0
    assert (img.double() - ref_img).abs().max() <= 1e-5 * max(1.0, float(ref_img.abs().max()))
    for name in ("means", "quats", "scales", "densities"):
        got, exp = getattr(grads, name), getattr(ref, name)
        assert got.shape == exp.shape
        if exp.numel() == 0:
            continue
        if case == "single_tile_grid":
            assert (got.double() - exp).abs().max() <= 1e-4 * float(exp.abs().max()), name
        else:  # nothing drawn: exact zeros on both paths
            assert torch.count_nonzero(got) == 0 and torch.count_nonzero(exp) == 0, name

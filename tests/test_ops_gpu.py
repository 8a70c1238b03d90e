"""Operator-level parity of the HIP library against the float64 CPU oracle.

Floating-point outputs: float64 kernels within 1e-10 relative, float32 within
the stated tolerances (outputs 2e-5, gradients 1e-4, max-abs relative to the
largest reference value). Integer outputs (radii, tile counts, keys, ids,
offsets) are compared bit-exactly.
"""

import json
import math
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import splat as O
from r2splat_amd import hip as ops

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
DTYPES = [torch.float32, torch.float64]


def tol(dtype, f32=2e-5, f64=1e-10):
    return f32 if dtype == torch.float32 else f64


def assert_rel(actual, expected, bound, what=""):
    err = O.relative_error(actual, expected)
    assert err <= bound, f"{what}: relative error {err:.3e} > {bound:.1e}"


def rand(g, *shape):
    return torch.rand(*shape, generator=g, dtype=torch.float64)


# --------------------------------------------------------------------------- parameter transform
@pytest.mark.parametrize("dtype", DTYPES)
def test_parameter_transform_fwd_bwd(device, dtype):
    g = torch.Generator().manual_seed(381)
    q = torch.randn(33, 4, generator=g, dtype=torch.float64) * torch.linspace(0.2, 4.0, 33, dtype=torch.float64)[:, None]
    s = 0.15 + rand(g, 33, 3)
    up = torch.randn(33, 6, generator=g, dtype=torch.float64)
    qr, sr = q.clone().requires_grad_(), s.clone().requires_grad_()
    ref = O.pack_sym(O.covariance(qr, sr))
    gq, gs = torch.autograd.grad((ref * up).sum(), (qr, sr))
    cov, mx = ops.parameter_transform_fwd(q.to(device, dtype), s.to(device, dtype))
    vq, vs = ops.parameter_transform_bwd(q.to(device, dtype), s.to(device, dtype), up.to(device, dtype))
    t = tol(dtype, 3e-6, 1e-12)
    assert_rel(cov, ref, t, "covars")
    assert torch.equal(mx.cpu().double(), s.to(dtype).max(dim=1).values.double())
    assert_rel(vq, gq, t * 10, "v_quats")
    assert_rel(vs, gs, t * 10, "v_scales")


@pytest.mark.parametrize("dtype,tiny,huge", [(torch.float32, 1e-20, 1e20), (torch.float64, 1e-200, 1e200)])
def test_parameter_transform_is_quaternion_scale_invariant(device, dtype, tiny, huge):
    q = torch.tensor([[0.7, -0.2, 0.4, 0.1]], dtype=torch.float64)
    s = torch.tensor([[0.23, 0.57, 1.2]], dtype=torch.float64)
    up = torch.tensor([[0.2, -0.4, 0.8, 0.6, -0.7, 0.3]], dtype=torch.float64)
    qr, sr = q.clone().requires_grad_(), s.clone().requires_grad_()
    ref = O.pack_sym(O.covariance(qr, sr))
    gq, gs = torch.autograd.grad((ref * up).sum(), (qr, sr))
    f = torch.tensor([1.0, -1.0, 0.125, -4.0, tiny, huge], dtype=torch.float64)[:, None]
    qs = (q * f).to(device, dtype)
    ss = s.expand(6, 3).to(device, dtype).contiguous()
    us = up.expand(6, 6).to(device, dtype).contiguous()
    cov, _ = ops.parameter_transform_fwd(qs, ss)
    vq, vs = ops.parameter_transform_bwd(qs, ss, us)
    assert bool(torch.isfinite(cov).all() and torch.isfinite(vq).all() and torch.isfinite(vs).all())
    t = tol(dtype, 3e-5, 3e-11)
    torch.testing.assert_close(cov.cpu().double(), ref.expand(6, 6), atol=t, rtol=t)
    torch.testing.assert_close(vq.cpu().double() * f, gq.expand(6, 4), atol=t, rtol=t)
    torch.testing.assert_close(vs.cpu().double(), gs.expand(6, 3), atol=t, rtol=t)
    # the adjoint of a scale-invariant map is tangent to the sphere
    torch.testing.assert_close((vq.cpu().double() * qs.cpu().double()).sum(-1), torch.zeros(6, dtype=torch.float64),
                               atol=t, rtol=0)


def test_parameter_transform_float32_subnormal_vjp(device):
    factor = 1e38
    q = torch.tensor([[0.7, -0.2, 0.4, 0.1]], dtype=torch.float32) * factor
    s = torch.tensor([[0.23, 0.57, 1.2]], dtype=torch.float32)
    up = torch.tensor([[0.2, -0.4, 0.8, 0.6, -0.7, 0.3]], dtype=torch.float32)
    qr = (q.double() / factor).requires_grad_()
    sr = s.double().requires_grad_()
    gq, gs = torch.autograd.grad((O.pack_sym(O.covariance(qr, sr)) * up.double()).sum(), (qr, sr))
    vq, vs = ops.parameter_transform_bwd(q.to(device), s.to(device), up.to(device))
    vq = vq.cpu().double()
    assert bool((vq.abs() > 0).all()) and bool((vq.abs() < torch.finfo(torch.float32).tiny).all())
    torch.testing.assert_close(vq * factor, gq, atol=3e-6, rtol=3e-5)
    torch.testing.assert_close(vs.cpu().double(), gs, atol=3e-6, rtol=3e-5)


# --------------------------------------------------------------------------- projection + conics
def _random_cameras(g, C, W, H, dtype=torch.float64):
    views = []
    for _ in range(C):
        R = torch.linalg.qr(torch.randn(3, 3, generator=g, dtype=torch.float64))[0]
        if torch.det(R) < 0:
            R[:, 0] = -R[:, 0]
        V = torch.eye(4, dtype=torch.float64)
        V[:3, :3] = R
        V[:3, 3] = torch.tensor([0.05, -0.1, 2.5], dtype=torch.float64)
        views.append(V)
    K = torch.tensor([[1.1 * W, 0, W / 2], [0, 1.1 * W, H / 2], [0, 0, 1]], dtype=torch.float64)
    return torch.stack(views), K.expand(C, 3, 3).contiguous()


@pytest.mark.parametrize("dtype", DTYPES)
def test_projection_fwd_bwd(device, dtype):
    g = torch.Generator().manual_seed(11)
    N, C, W, H = 200, 3, 40, 30
    means = (rand(g, N, 3) - 0.5) * 1.6
    cov = O.pack_sym(O.covariance(torch.randn(N, 4, generator=g, dtype=torch.float64), 0.02 + 0.1 * rand(g, N, 3)))
    views, Ks = _random_cameras(g, C, W, H)
    d = lambda t: t.to(device, dtype).contiguous()
    m2, c2, radii = ops.world_to_image_projection_fwd(d(means), d(cov), d(views), d(Ks), W, H, 1e-3, 3.0)
    mr, cr = means.clone().requires_grad_(), cov.clone().requires_grad_()
    rm2, rc2, vis = O.project(mr, O.unpack_sym(cr), views, Ks, W, H)
    rc2p = O.pack_sym(rc2)
    assert_rel(m2, rm2, tol(dtype, 1e-6), "means2d")
    assert_rel(c2, rc2p, tol(dtype, 1e-5), "cov2d")
    # radii are integers computed from the kernel's own cov2d/means2d: bit-exact
    expect = O.radii_2d(c2.cpu().numpy(), m2.cpu().numpy(), 1e-3, 3.0, W, H)
    assert np.array_equal(radii.cpu().numpy(), expect)
    v_m2 = torch.randn(C, N, 2, generator=g, dtype=torch.float64)
    v_c2 = torch.randn(C, N, 6, generator=g, dtype=torch.float64)
    gm, gc = torch.autograd.grad((rm2 * v_m2).sum() + (rc2p * v_c2).sum(), (mr, cr))
    vm, vc = ops.world_to_image_projection_bwd(d(means), d(cov), d(views), d(Ks), W, H, d(v_m2), d(v_c2))
    assert_rel(vm, gm, tol(dtype, 1e-5), "v_means")
    assert_rel(vc, gc, tol(dtype, 1e-5), "v_covars")


def test_projection_culls_points_behind_camera(device):
    means = torch.tensor([[0.0, 0.0, 0.005], [0.0, 0.0, -1.0], [0.0, 0.0, 2.0]], device=device)
    cov = torch.tensor([[0.01, 0, 0, 0.01, 0, 0.01]], device=device).expand(3, 6).contiguous()
    view = torch.eye(4, device=device)[None]
    K = torch.tensor([[[10.0, 0, 5], [0, 10.0, 5], [0, 0, 1]]], device=device)
    m2, c2, radii = ops.world_to_image_projection_fwd(means, cov, view, K, 10, 10, 0.0, 3.0)
    assert float(m2[0, :2].abs().sum()) == 0 and float(c2[0, :2].abs().sum()) == 0
    assert radii[0, :2].abs().sum().item() == 0 and radii[0, 2].min().item() > 0


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("scale", [0.5, 0.01, 1e-5])
@pytest.mark.parametrize("eps", [0.0, 1e-4, 0.3])
def test_conics_mu(device, dtype, scale, eps):
    cov = torch.tensor([[[1.2, 0.2, -0.1, 0.9, 0.15, 1.5], [0.7, -0.12, 0.16, 1.4, -0.08, 1.1]]],
                       dtype=torch.float64) * scale ** 2
    cov_d = cov.to(device, dtype)
    conics, mu = ops.conics_mu_fwd(cov_d, eps)
    cr = cov_d.double().cpu().requires_grad_()
    rcon, rmu = O.conics_mu(O.unpack_sym(cr), eps)
    t = tol(dtype, 3e-5, 1e-10)
    torch.testing.assert_close(conics.cpu().double(), rcon, atol=t * max(1.0, rcon.abs().max().item()), rtol=t)
    torch.testing.assert_close(mu.cpu().double(), rmu, atol=0, rtol=t)
    for source in ("conics", "mu"):
        vc = torch.zeros_like(conics)
        vm = torch.zeros_like(mu)
        if source == "conics":
            vc[:] = torch.tensor([0.3, -0.7, 1.1], dtype=dtype, device=device) * (scale ** 2 + eps) ** 2
        else:
            vm[:] = torch.tensor([[0.8, -1.3]], dtype=dtype, device=device)
        got = ops.conics_mu_bwd(cov_d, eps, conics, mu, vc, vm)
        exp = torch.autograd.grad((rcon * vc.cpu().double()).sum() + (rmu * vm.cpu().double()).sum(), cr,
                                  retain_graph=True)[0]
        torch.testing.assert_close(got.cpu().double(), exp, atol=t * max(1.0, exp.abs().max().item()), rtol=t)


def test_conics_mu_degenerate_gives_zero_outputs_and_gradients(device):
    cov = torch.tensor([[[0, 0, 0, 0, 0, 0], [1, 1, 0, 1, 0, 1]]], dtype=torch.float64, device=device)
    conics, mu = ops.conics_mu_fwd(cov, 0)
    grad = ops.conics_mu_bwd(cov, 0, conics, mu, torch.ones_like(conics), torch.ones_like(mu))
    assert conics.abs().sum().item() == 0 and mu.abs().sum().item() == 0 and grad.abs().sum().item() == 0


# --------------------------------------------------------------------------- binning
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("tile", [16, 15, 4])
def test_isect_tiles_2d_bit_exact(device, dtype, tile):
    rng = np.random.default_rng(tile)
    C, N, W, H = 3, 500, 70, 45
    ntx, nty = -(-W // tile), -(-H // tile)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    m2 = rng.uniform(-10, [W + 10, H + 10], size=(C, N, 2)).astype(npdt)
    radii = rng.integers(-1, 14, size=(C, N, 2)).astype(np.int32)
    counts, ids, flat = ops.isect_tiles(torch.from_numpy(m2).to(device), torch.from_numpy(radii).to(device), tile,
                                        ntx, nty)
    rc, rids, rflat = O.isect_tiles_2d(m2, radii, tile, ntx, nty)
    assert np.array_equal(counts.cpu().numpy(), rc)
    assert np.array_equal(ids.cpu().numpy(), rids)
    assert np.array_equal(flat.cpu().numpy(), rflat)
    off = ops.isect_offset_encode(ids, C, ntx, nty)
    assert tuple(off.shape) == (C, nty, ntx)
    assert np.array_equal(off.cpu().numpy().reshape(-1), O.offsets(rids, C * ntx * nty))


def test_isect_tiles_empty_and_square_radii(device):
    m2 = torch.zeros(2, 3, 2, device=device)
    radii = torch.zeros(2, 3, dtype=torch.int32, device=device)
    counts, ids, flat = ops.isect_tiles(m2, radii, 16, 2, 2)
    assert ids.numel() == 0 and flat.numel() == 0 and counts.abs().sum().item() == 0
    off = ops.isect_offset_encode(ids, 2, 2, 2)
    assert off.abs().sum().item() == 0
    radii = torch.full((2, 3), 5, dtype=torch.int32, device=device)
    m2[..., 0] = 8.0
    m2[..., 1] = 20.0
    counts, ids, flat = ops.isect_tiles(m2, radii, 16, 2, 2)
    rc, rids, rflat = O.isect_tiles_2d(m2.cpu().numpy(), radii.cpu().numpy()[..., None].repeat(2, -1), 16, 2, 2)
    assert np.array_equal(ids.cpu().numpy(), rids) and np.array_equal(flat.cpu().numpy(), rflat)


@pytest.mark.parametrize("dtype", DTYPES)
def test_isect_tiles_3d_known_answers(device, dtype):
    kat = json.load(open(GOLDEN / "tile_bounds_3d.json"))
    nx, ny, nz = kat["n_tiles_xyz"]
    for case in kat["cases"]:
        means = torch.tensor([case["mean"]], dtype=dtype, device=device)
        radii = torch.full((1, 3), case["radius"], dtype=torch.int32, device=device)
        counts, ids, flat = ops.isect_tiles_3d(means, radii, kat["tile_size"], nx, ny, nz)
        assert counts.tolist() == [len(case["expected_ids"])], case["id"]
        assert ids.tolist() == case["expected_ids"], case["id"]
        assert flat.tolist() == [0] * len(case["expected_ids"])
        off = ops.isect_offset_encode_3d(ids, nx, ny, nz)
        assert tuple(off.shape) == (nz, ny, nx)
        expect = np.searchsorted(np.array(case["expected_ids"], dtype=np.int64), np.arange(24))
        assert off.reshape(-1).tolist() == expect.tolist()


def test_isect_tiles_3d_random_bit_exact(device):
    rng = np.random.default_rng(5)
    N, T, nx, ny, nz = 3000, 8, 9, 7, 5
    m = rng.uniform(-10, [nx * T + 10, ny * T + 10, nz * T + 10], size=(N, 3)).astype(np.float32)
    r = rng.integers(0, 12, size=(N, 3)).astype(np.int32)
    counts, ids, flat = ops.isect_tiles_3d(torch.from_numpy(m).to(device), torch.from_numpy(r).to(device), T, nx, ny, nz)
    rc, rids, rflat = O.isect_tiles_3d(m, r, T, nx, ny, nz)
    assert np.array_equal(counts.cpu().numpy(), rc)
    assert np.array_equal(ids.cpu().numpy(), rids)
    assert np.array_equal(flat.cpu().numpy(), rflat)


def test_radix_sort_many_blocks_and_passes(device):
    # 3 sort passes (key width 21 bits) over several thousand tiles, many equal keys.
    rng = np.random.default_rng(9)
    C, N, T, W, H = 40, 6000, 4, 256, 200
    ntx, nty = W // T, H // T
    m2 = rng.uniform(0, [W, H], size=(C, N, 2)).astype(np.float32)
    radii = rng.integers(1, 6, size=(C, N, 2)).astype(np.int32)
    counts, ids, flat = ops.isect_tiles(torch.from_numpy(m2).to(device), torch.from_numpy(radii).to(device), T,
                                        ntx, nty)
    rc, rids, rflat = O.isect_tiles_2d(m2, radii, T, ntx, nty)
    assert ids.numel() == rids.size and ids.numel() > 10 * 4096
    assert np.array_equal(ids.cpu().numpy(), rids)
    assert np.array_equal(flat.cpu().numpy(), rflat)


# --------------------------------------------------------------------------- rasterization (low level)
def _raster_scene(g, C=2, N=24, W=37, H=29, beta=False):
    m2 = rand(g, C, N, 2) * torch.tensor([W, H], dtype=torch.float64)
    A = torch.randn(C, N, 2, 2, generator=g, dtype=torch.float64)
    P = (A @ A.transpose(-1, -2)) * 0.05 + 0.02 * torch.eye(2, dtype=torch.float64)
    conics = P[..., [0, 0, 1], [0, 1, 1]].contiguous()
    opac = 0.2 + rand(g, C, N)
    mus = 0.5 + rand(g, C, N)
    b2 = (1.0 + 3 * rand(g, C, N)) if beta else None
    unc = 0.3 * rand(g, N)
    return m2, conics, opac, mus, b2, unc


def _bin(m2, tile, W, H, device):
    """Bins with a generous radius so every pair inside a tile range is exercised."""
    C, N = m2.shape[:2]
    radii = torch.full((C, N, 2), 12, dtype=torch.int32)
    ntx, nty = -(-W // tile), -(-H // tile)
    _, ids, flat = ops.isect_tiles(m2.to(device), radii.to(device), tile, ntx, nty)
    off = ops.isect_offset_encode(ids, C, ntx, nty)
    mask = O.tile_mask_2d(m2.numpy(), radii.numpy(), tile, W, H)
    return radii, ids, flat, off, mask


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("beta", [False, True])
@pytest.mark.parametrize("tile", [16, 7])
def test_rasterize_ops_forward_and_all_vjps(device, dtype, beta, tile):
    g = torch.Generator().manual_seed(17 + tile)
    W, H = 37, 29
    m2, conics, opac, mus, b2, unc = _raster_scene(g, W=W, H=H, beta=beta)
    _, ids, flat, off, mask = _bin(m2, tile, W, H, device)
    d = lambda t: None if t is None else t.to(device, dtype).contiguous()
    leaves = [t.clone().requires_grad_() for t in (m2, conics, opac, mus)] + ([b2.clone().requires_grad_()] if beta else [])
    ur = unc.clone().requires_grad_()
    ref, ref_var = O.rasterize(*leaves[:4], W, H, beta2d=leaves[4] if beta else None, unc=ur, mask=mask)
    args = [d(m2), d(conics), d(opac), d(mus)] + ([d(b2)] if beta else [])
    prefix = "beta_" if beta else ""
    t = tol(dtype)
    img = getattr(ops, prefix + "rasterize_to_pixels_fwd")(*args, off, flat, W, H, tile)
    assert_rel(img, ref, t, "image")
    img2, var = getattr(ops, prefix + "rasterize_to_pixels_mean_var_fwd")(*args, d(unc), off, flat, W, H, tile)
    assert_rel(img2, ref, t, "image(mean_var)")
    assert_rel(var, ref_var, t, "variance")
    v = torch.linspace(-0.4, 1.3, ref.numel(), dtype=torch.float64).reshape(ref.shape)
    vv = torch.linspace(0.9, -0.2, ref.numel(), dtype=torch.float64).reshape(ref.shape)
    exp_mean = torch.autograd.grad((ref * v).sum(), leaves, retain_graph=True)
    exp_abs = O.abs_mean2d_grad(m2, conics, opac, mus, W, H, v, beta2d=b2, mask=mask)
    gt = tol(dtype, 1e-4, 1e-9)
    for name, fn, extra in (("pixel", prefix + "rasterize_to_pixels_bwd", (off, flat)),
                            ("primitive", prefix + "rasterize_to_pixels_bwd_per_gaussian", (off, ids, flat))):
        grads = getattr(ops, fn)(*args, *extra, d(v), W, H, tile, True)
        assert grads[0] is not None and len(grads) == 5 + int(beta)
        for k, (a, e) in enumerate(zip(grads[1:], exp_mean)):
            assert_rel(a, e, gt, f"{name} grad {k}")
        # absgrad: sum over pixels of |d(v alpha)/d mean2d| (RasterizeToPixelsBwd.cu:219-221, 250-255)
        assert_rel(grads[0], exp_abs, gt, f"{name} absgrad")
    exp_joint = torch.autograd.grad((ref * v + ref_var * vv).sum(), leaves + [ur], retain_graph=True)
    fused = getattr(ops, prefix + "rasterize_to_pixels_mean_var_bwd_per_gaussian")(
        *args, off, ids, flat, d(v), d(vv), d(unc), False, W, H, tile, False)
    assert fused[0] is None and len(fused) == 6 + int(beta)
    for k, (a, e) in enumerate(zip(fused[1:], exp_joint)):
        assert_rel(a, e, gt, f"fused grad {k}")
    with_abs = getattr(ops, prefix + "rasterize_to_pixels_mean_var_bwd_per_gaussian")(
        *args, off, ids, flat, d(v), d(vv), d(unc), False, W, H, tile, True)
    exp_abs_var = O.abs_mean2d_grad(m2, conics, opac, mus, W, H, v, beta2d=b2, unc=unc, v_var=vv, mask=mask)
    assert_rel(with_abs[0], exp_abs_var, gt, "mean_var absgrad")
    vq = getattr(ops, prefix + "rasterize_to_pixels_uncertainty_bwd")(*args, off, flat, d(vv), W, H, tile)
    assert_rel(vq, exp_joint[-1], gt, "uncertainty bwd")
    # detached overload: mean adjoint only from v, uncertainty gradient unchanged
    det = getattr(ops, prefix + "rasterize_to_pixels_mean_var_bwd_per_gaussian")(
        *args, off, ids, flat, d(v), d(vv), W, H, tile, False)
    for k, (a, e) in enumerate(zip(det[1:-1], exp_mean)):
        assert_rel(a, e, gt, f"detached grad {k}")
    assert_rel(det[-1], exp_joint[-1], gt, "detached uncertainty")


def test_rasterize_single_tile_closed_form(device):
    means2d = torch.tensor([[[2.5, 2.5]]], device=device)
    conics = torch.tensor([[[1.0, 0.1, 1.2]]], device=device)
    opac = torch.tensor([[2.0]], device=device)
    mu = torch.tensor([[3.0]], device=device)
    off = torch.zeros((1, 1, 1), dtype=torch.int32, device=device)
    flat = torch.zeros((1,), dtype=torch.int32, device=device)
    out = ops.rasterize_to_pixels_fwd(means2d, conics, opac, mu, off, flat, 5, 5, 8)
    yy, xx = torch.meshgrid(torch.arange(5.0) + 0.5, torch.arange(5.0) + 0.5, indexing="ij")
    dx, dy = xx - 2.5, yy - 2.5
    s = 0.5 * (dx * dx + 1.2 * dy * dy) + 0.1 * dx * dy
    ref = 6.0 * torch.exp(-s)
    ref = torch.where(ref >= 1e-5, ref, 0.0)
    torch.testing.assert_close(out.cpu()[0], ref, atol=1e-6, rtol=1e-6)


# --------------------------------------------------------------------------- volume mapping + voxels
@pytest.mark.parametrize("dtype", DTYPES)
def test_volume_mapping_fwd_bwd(device, dtype):
    g = torch.Generator().manual_seed(21)
    N = 40
    means = rand(g, N, 3) - 0.5
    cov = O.pack_sym(O.covariance(torch.randn(N, 4, generator=g, dtype=torch.float64), 0.02 + 0.2 * rand(g, N, 3)))
    ext = torch.tensor([2.0, 3.0, 4.0], dtype=torch.float64)
    cen = torch.tensor([0.2, -0.3, 1.0], dtype=torch.float64)
    shape = (5, 9, 13)  # z, y, x
    d = lambda t: t.to(device, dtype).contiguous()
    mapped, conics, radii = ops.world_to_volume_mapping_fwd(d(means), d(cov), d(torch.ones(N)), d(ext), d(cen),
                                                            13, 9, 5, 1e-3, 3.0)
    mr, cr = means.clone().requires_grad_(), cov.clone().requires_grad_()
    rm, P, ok = O.volume_mapping(mr, O.unpack_sym(cr), ext, cen, shape, 1e-3)
    assert bool(ok.all())
    t = tol(dtype, 3e-6, 1e-11)
    assert_rel(mapped, rm, t, "mapped")
    assert_rel(conics, O.pack_sym(P), t, "conics")
    r = torch.tensor([13.0, 9.0, 5.0], dtype=torch.float64) / ext
    sv = O.unpack_sym(cov) * r[:, None] * r[None, :] + 1e-3 * torch.eye(3, dtype=torch.float64)
    expect = torch.ceil(3.0 * sv.diagonal(dim1=-2, dim2=-1).sqrt()).to(torch.int32)
    assert torch.equal(radii.cpu(), expect) or dtype == torch.float32  # float32 may round across an integer
    vm = torch.randn(N, 3, generator=g, dtype=torch.float64)
    vc = torch.randn(N, 6, generator=g, dtype=torch.float64)
    gm, gc = torch.autograd.grad((rm * vm).sum() + (O.pack_sym(P) * vc).sum(), (mr, cr))
    am, ac = ops.world_to_volume_mapping_bwd(d(means), d(cov), d(ext), d(cen), 13, 9, 5, 1e-3, d(vm), d(vc))
    assert_rel(am, gm, t, "v_means")
    assert_rel(ac, gc, tol(dtype, 1e-5, 1e-10), "v_covars")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("scale", [1e-6, 1.0, 1e6])
def test_volume_mapping_invalid_covariances_are_zero(device, dtype, scale):
    covars = torch.tensor([[0, 0, 0, 0, 0, 0], [1, 0, 0, 1, 0, 0], [1, 1, 1, 1, 1, 1], [1, 2, 0, 1, 0, 1],
                           [-1, 0, 0, -1, 0, 1]], dtype=dtype, device=device) * scale
    means = covars.new_zeros((5, 3))
    one3, zero3 = covars.new_ones(3), covars.new_zeros(3)
    _, conics, radii = ops.world_to_volume_mapping_fwd(means, covars, covars.new_ones(5), one3, zero3, 1, 1, 1, 0.0, 3.0)
    assert conics.abs().sum().item() == 0 and radii.abs().sum().item() == 0
    _, grad = ops.world_to_volume_mapping_bwd(means, covars, one3, zero3, 1, 1, 1, 0.0, torch.zeros_like(means),
                                              torch.ones_like(conics))
    assert grad.abs().sum().item() == 0


def _vox_scene(g, N=20, beta=False):
    m = rand(g, N, 3) * torch.tensor([9.0, 7.0, 6.0], dtype=torch.float64)
    A = torch.randn(N, 3, 3, generator=g, dtype=torch.float64)
    P = (A @ A.transpose(-1, -2)) * 0.08 + 0.05 * torch.eye(3, dtype=torch.float64)
    conics = O.pack_sym(P)
    opac = 0.3 + rand(g, N)
    b = (1.0 + 3 * rand(g, N)) if beta else None
    unc = 0.3 * rand(g, N)
    return m, conics, opac, b, unc


def _vox_dense(m, conics, opac, beta, unc, shape, mask, detach=False):
    """Voxel-frame dense reference (tile-masked)."""
    D, H, W = shape
    zz, yy, xx = torch.meshgrid(torch.arange(D, dtype=torch.float64) + 0.5, torch.arange(H, dtype=torch.float64) + 0.5,
                                torch.arange(W, dtype=torch.float64) + 0.5, indexing="ij")
    grid = torch.stack([xx, yy, zz], -1)
    P = O.unpack_sym(conics)
    mean = torch.zeros(shape, dtype=torch.float64)
    var = torch.zeros(shape, dtype=torch.float64)
    for n in range(m.shape[0]):
        delta = grid - m[n]
        d2 = torch.einsum("...i,ij,...j->...", delta, P[n], delta)
        if beta is None:
            a = opac[n] * torch.exp(-0.5 * d2)
            keep = (d2 >= 0) & (a.abs() >= 1e-6)
        else:
            inside = (d2 >= 0) & (d2 < 1)
            h = torch.where(inside, 1 - d2, torch.ones_like(d2))
            a = torch.where(inside, opac[n] * h.pow(beta[n]), torch.zeros_like(d2))
            keep = inside & (a.abs() >= 1e-5)
        keep = keep & torch.as_tensor(mask[n])
        a = torch.where(keep, a, torch.zeros_like(a))
        mean = mean + a
        var = var + unc[n] * (a.detach() if detach else a).square()
    return mean, var


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("beta", [False, True])
@pytest.mark.parametrize("tile", [8, 3])
def test_voxelize_ops_forward_and_vjps(device, dtype, beta, tile):
    g = torch.Generator().manual_seed(31 + tile)
    shape = (6, 7, 9)
    D, H, W = shape
    m, conics, opac, b, unc = _vox_scene(g, beta=beta)
    radii = torch.full((m.shape[0], 3), 5, dtype=torch.int32)
    nx, ny, nz = -(-W // tile), -(-H // tile), -(-D // tile)
    _, ids, flat = ops.isect_tiles_3d(m.to(device), radii.to(device), tile, nx, ny, nz)
    off = ops.isect_offset_encode_3d(ids, nx, ny, nz)
    mask = O.tile_mask_3d(m.numpy(), radii.numpy(), tile, shape)
    leaves = [t.clone().requires_grad_() for t in (m, conics, opac)] + ([b.clone().requires_grad_()] if beta else [])
    ur = unc.clone().requires_grad_()
    ref, ref_var = _vox_dense(*leaves[:3], leaves[3] if beta else None, ur, shape, mask)
    d = lambda t: t.to(device, dtype).contiguous()
    args = [d(m), d(conics), d(opac)] + ([d(b)] if beta else [])
    prefix = "beta_" if beta else ""
    t = tol(dtype)
    out = getattr(ops, prefix + "voxelize_to_voxels_fwd")(*args, off, flat, W, H, D, tile)
    assert_rel(out, ref, t, "volume")
    out2, var = getattr(ops, prefix + "voxelize_to_voxels_mean_var_fwd")(*args, d(unc), off, flat, W, H, D, tile)
    assert_rel(out2, ref, t, "volume(mean_var)")
    assert_rel(var, ref_var, t, "variance")
    v = torch.linspace(-0.4, 1.3, ref.numel(), dtype=torch.float64).reshape(shape)
    vv = torch.linspace(0.9, -0.2, ref.numel(), dtype=torch.float64).reshape(shape)
    gt = tol(dtype, 1e-4, 1e-9)
    exp_mean = torch.autograd.grad((ref * v).sum(), leaves, retain_graph=True)
    for name, fn, extra in (("tile", prefix + "voxelize_to_voxels_bwd", (off, flat)),
                            ("primitive", prefix + "voxelize_to_voxels_bwd_per_gaussian", (off, ids, flat))):
        grads = getattr(ops, fn)(*args, *extra, d(v), W, H, D, tile)
        assert len(grads) == 3 + int(beta)
        for k, (a, e) in enumerate(zip(grads, exp_mean)):
            assert_rel(a, e, gt, f"{name} grad {k}")
    exp_joint = torch.autograd.grad((ref * v + ref_var * vv).sum(), leaves + [ur])
    fused = getattr(ops, prefix + "voxelize_to_voxels_mean_var_bwd_per_gaussian")(
        *args, off, ids, flat, d(v), d(vv), d(unc), False, W, H, D, tile)
    assert len(fused) == 4 + int(beta)
    for k, (a, e) in enumerate(zip(fused, exp_joint)):
        assert_rel(a, e, gt, f"fused grad {k}")
    det = getattr(ops, prefix + "voxelize_to_voxels_mean_var_bwd_per_gaussian")(
        *args, off, ids, flat, d(v), d(vv), W, H, D, tile)
    for k, (a, e) in enumerate(zip(det[:-1], exp_mean)):
        assert_rel(a, e, gt, f"detached grad {k}")
    assert_rel(det[-1], exp_joint[-1], gt, "detached uncertainty")


# --------------------------------------------------------------------------- losses
def _f32(x):
    return float(np.float32(x))


@pytest.mark.parametrize("dtype", DTYPES)
def test_volume_tv_loss(device, dtype):
    """lambda is a float in the reference's op (Ops.h:36-44): float64 volumes are
    scaled by float32(0.1) = 0.100000001490116..., not by 0.1."""
    g = torch.Generator().manual_seed(0)
    vol = torch.randn(2, 4, 5, 6, generator=g, dtype=torch.float64)
    vr = vol.clone().requires_grad_()
    tv = O.volume_tv(vr)
    lam = _f32(0.1)
    (grad,) = torch.autograd.grad(lam * tv, vr)
    loss, tv_d, tv_map = ops.volume_loss_fwd(vol.to(device, dtype), 0.1)
    t = tol(dtype, 1e-6, 1e-13)
    assert abs(tv_d.item() - tv.item()) <= t * abs(tv.item())
    assert abs(loss.item() - lam * tv.item()) <= t * abs(tv.item())
    gd = ops.volume_loss_bwd(vol.to(device, dtype), torch.ones(1, device=device, dtype=dtype), 0.1)
    torch.testing.assert_close(gd.cpu().double(), grad, atol=t, rtol=t)
    if dtype == torch.float64:
        # the float rounding of lambda is visible at float64 precision
        assert abs(loss.item() - 0.1 * tv.item()) > 1e-9 * abs(tv.item())


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(2, 4, 5, 6), (1, 1, 1, 7), (1, 33, 17, 65)])
def test_volume_tv_iso_loss(device, dtype, shape):
    g = torch.Generator().manual_seed(1)
    vol = torch.randn(*shape, generator=g, dtype=torch.float64)
    vr = vol.clone().requires_grad_()
    tv = O.volume_tv_iso(vr, 1e-6)
    (grad,) = torch.autograd.grad(1e-3 * tv, vr)
    loss, tv_d, tv_map = ops.volume_loss_iso_fwd(vol.to(device, dtype), 1e-3, eps=1e-6)
    t = tol(dtype, 2e-6, 1e-12)
    assert abs(tv_d.item() - tv.item()) <= t * abs(tv.item())
    assert abs(loss.item() - 1e-3 * tv.item()) <= t * abs(tv.item())
    assert abs(float(tv_map.double().sum().cpu()) / vol.numel() - tv.item()) <= t * abs(tv.item())
    gd = ops.volume_loss_iso_bwd(vol.to(device, dtype), torch.ones(1, device=device, dtype=dtype), 1e-3, eps=1e-6)
    assert O.relative_error(gd, grad) <= tol(dtype, 1e-5, 1e-12)


def test_image_l1_loss(device):
    p = torch.tensor([[[-2.0, 1.0], [3.0, -4.0]]], device=device)
    t = torch.zeros_like(p)
    loss, l1, _ = ops.image_loss_fwd(p, t, 0.25)
    assert abs(loss.item() - 0.25 * 2.5) < 1e-6
    g = ops.image_loss_bwd(p, t, torch.ones(1, device=device), 0.25)
    torch.testing.assert_close(g.cpu(), 0.25 * torch.sign(p.cpu()) / 4)
    # float64: lambda is rounded to float first, as the reference's float argument is
    pd, td = p.double(), t.double()
    loss, _, _ = ops.image_loss_fwd(pd, td, 0.1)
    assert loss.item() == _f32(0.1) * 2.5
    g = ops.image_loss_bwd(pd, td, torch.ones(1, device=device, dtype=torch.float64), 0.1)
    assert torch.equal(g.cpu(), torch.sign(pd.cpu()) * _f32(0.1) / 4)


def test_ops_reject_bad_inputs_with_runtime_error(device):
    q = torch.randn(2, 4, device=device, dtype=torch.float64)
    s = torch.rand(2, 3, device=device, dtype=torch.float64)
    with pytest.raises(RuntimeError):
        ops.parameter_transform_fwd(q[:, :3].contiguous(), s)
    with pytest.raises(RuntimeError):
        ops.parameter_transform_fwd(q, s.float())
    with pytest.raises(RuntimeError):
        ops.parameter_transform_bwd(q, s, torch.zeros(2, 5, device=device, dtype=torch.float64))
    with pytest.raises(RuntimeError):
        ops.parameter_transform_fwd(q.cpu(), s.cpu())

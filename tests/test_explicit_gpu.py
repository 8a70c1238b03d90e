"""Public drop-in API on the GPU: parity with the float64 oracle, finite
differences, and the (output, ctx) / ParameterGradients contracts."""

from dataclasses import is_dataclass

import numpy as np
import pytest
import torch

import r2splat_amd
from oracle import splat as O
from r2splat_amd import hip as ops
from r2splat_amd import workloads

pytestmark = pytest.mark.gpu
FIELDS = ("means", "quats", "scales", "densities", "beta", "uncertainty")


def _scene(device, use_beta=True, uncertainty=True, dtype=torch.float64, requires_grad=False):
    def t(data):
        return torch.tensor(data, device=device, dtype=dtype, requires_grad=requires_grad)

    values = {
        "means": t([[-0.12, 0.09, 1.94], [0.18, -0.13, 2.12]]),
        "quats": t([[1.7, 0.2, -0.15, 0.08], [0.8, -0.12, 0.16, 0.11]]),
        "scales": t([[0.32, 0.43, 0.37], [0.41, 0.34, 0.48]]),
        "densities": t([1.2, 0.8]),
        "beta": t([2.2, 5.5]) if use_beta else None,
        "uncertainty": t([0.17, 0.29]) if uncertainty else None,
    }
    viewmats = torch.eye(4, device=device, dtype=dtype).repeat(2, 1, 1)
    viewmats[1, :2, 3] = viewmats.new_tensor([0.12, -0.05])
    values.update(viewmats=viewmats,
                  Ks=viewmats.new_tensor([[[12, 0, 4.3], [0, 10, 3.2], [0, 0, 1]]]).repeat(2, 1, 1),
                  volume_extent=viewmats.new_tensor([1.8, 1.6, 1.4]),
                  volume_center=viewmats.new_tensor([0, 0, 2]))
    return values


def _forward(op, values, *, mode="primitive", detach=False, eps=0.01):
    kwargs = dict(beta=values["beta"], uncertainty=values["uncertainty"], eps=eps, detach_variance=detach,
                  backward_mode=mode, radius_clip=1 if values["beta"] is not None else 3)
    inputs = [values[n] for n in FIELDS[:4]]
    if op == "rasterize":
        return r2splat_amd.rasterize_fwd(*inputs, values["viewmats"], values["Ks"], (7, 9), **kwargs)
    return r2splat_amd.voxelize_fwd(*inputs, (5, 7, 9), volume_extent=values["volume_extent"],
                                    volume_center=values["volume_center"], **kwargs)


def _cpu_leaves(values):
    return {k: (None if v is None else v.detach().cpu().double().requires_grad_(k in FIELDS))
            for k, v in values.items()}


def _oracle(op, ref, detach, eps=0.01, radius_clip=None, mask=None, shape=None):
    inputs = [ref[n] for n in FIELDS[:4]]
    if op == "rasterize":
        return O.rasterize_scene(*inputs, ref["viewmats"], ref["Ks"], shape or (7, 9), beta=ref["beta"],
                                 uncertainty=ref["uncertainty"], eps=eps, detach=detach, mask=mask)
    return O.voxelize(*inputs, shape or (5, 7, 9), ref["volume_extent"], ref["volume_center"], beta=ref["beta"],
                      uncertainty=ref["uncertainty"], eps=eps, detach=detach, mask=mask)


def _outputs(out):
    return out if isinstance(out, tuple) else (out,)


def _weights(out):
    res = []
    for k, v in enumerate(_outputs(out)):
        res.append(torch.linspace(-0.3 + k, 1.1 + k, v.numel(), device=v.device, dtype=v.dtype).reshape_as(v))
    return res


def _assert_grads(actual, expected, dtype):
    assert is_dataclass(actual)
    atol, rtol = (2e-4, 3e-4) if dtype == torch.float32 else (2e-9, 2e-9)
    for name in FIELDS:
        value = getattr(actual, name)
        if name not in expected:
            assert value is None
            continue
        assert value.grad_fn is None and not value.requires_grad
        assert bool(torch.isfinite(value).all())
        torch.testing.assert_close(value.cpu().double(), expected[name].to(torch.float64), atol=atol, rtol=rtol,
                                   msg=name)


@pytest.mark.parametrize("op", ["rasterize", "voxelize"])
@pytest.mark.parametrize("use_beta,uncertainty,detach,dtype,mode", [
    (False, False, False, torch.float32, "primitive"),
    (True, False, False, torch.float64, "pixel"),
    (False, False, False, torch.float64, "pixel"),
    (True, False, False, torch.float32, "primitive"),
    (False, True, False, torch.float64, "pixel"),
    (True, True, False, torch.float32, "primitive"),
    (False, True, True, torch.float32, "primitive"),
    (True, True, True, torch.float64, "pixel"),
])
def test_explicit_vjp_matches_oracle(device, op, use_beta, uncertainty, detach, dtype, mode):
    values = _scene(device, use_beta, uncertainty, dtype, requires_grad=True)
    for n in FIELDS:
        if values[n] is not None:
            values[n].grad = torch.full_like(values[n], 1.25)
    ref = _cpu_leaves(values)
    out, ctx = _forward(op, values, mode=mode, detach=detach)
    expected = _oracle(op, ref, detach)
    tol = 2e-5 if dtype == torch.float32 else 2e-11
    for a, e in zip(_outputs(out), _outputs(expected)):
        assert a.grad_fn is None and not a.requires_grad
        torch.testing.assert_close(a.cpu().double(), e.detach(), atol=tol, rtol=tol)
    weights = _weights(out)
    names = [n for n in FIELDS if ref[n] is not None]
    loss = sum((e * w.detach().cpu().double()).sum() for e, w in zip(_outputs(expected), weights))
    exp_grads = dict(zip(names, torch.autograd.grad(loss, [ref[n] for n in names])))
    got = getattr(r2splat_amd, op + "_bwd")(ctx, *weights)
    _assert_grads(got, exp_grads, dtype)
    for n in FIELDS:
        if values[n] is not None:
            torch.testing.assert_close(values[n].grad, torch.full_like(values[n], 1.25))


@pytest.mark.parametrize("op", ["rasterize", "voxelize"])
def test_explicit_vjp_matches_central_differences(device, op):
    values = _scene(device)
    out, ctx = _forward(op, values)
    weights = _weights(out)
    grads = getattr(r2splat_amd, op + "_bwd")(ctx, *weights)
    entries = {"means": (0, 0), "quats": (0, 1), "scales": (0, 2), "densities": (0,), "beta": (0,),
               "uncertainty": (0,)}
    step = 1e-4
    for name, index in entries.items():
        losses = []
        for sign in (1, -1):
            pert = {k: (None if v is None else v.clone()) for k, v in values.items()}
            pert[name][index] += sign * step
            sampled, _ = _forward(op, pert)
            losses.append(sum((s * w).sum() for s, w in zip(_outputs(sampled), weights)))
        fd = (losses[0] - losses[1]) / (2 * step)
        analytic = getattr(grads, name)[index]
        assert abs(analytic.item()) > 1e-9, name
        torch.testing.assert_close(analytic, fd, atol=2e-4, rtol=3e-3, msg=name)


@pytest.mark.parametrize("op", ["rasterize", "voxelize"])
@pytest.mark.parametrize("case,detach", [("zero", False), ("mean", False), ("variance", False), ("variance", True)])
def test_optional_adjoints_and_detached_variance(device, op, case, detach):
    values = _scene(device)
    out, ctx = _forward(op, values, detach=detach)
    v_mean, v_var = _weights(out)
    if case == "zero":
        v_mean = v_var = None
    elif case == "mean":
        v_var = None
    else:
        v_mean = None
    got = getattr(r2splat_amd, op + "_bwd")(ctx, v_mean, v_var)
    if case == "zero":
        for n in FIELDS:
            torch.testing.assert_close(getattr(got, n), torch.zeros_like(values[n]))
        return
    ref = _cpu_leaves(values)
    expected = _oracle(op, ref, detach)
    w = [torch.zeros_like(expected[0]) if v_mean is None else v_mean.cpu().double(),
         torch.zeros_like(expected[1]) if v_var is None else v_var.cpu().double()]
    names = list(FIELDS)
    loss = (expected[0] * w[0]).sum() + (expected[1] * w[1]).sum()
    exp_grads = dict(zip(names, torch.autograd.grad(loss, [ref[n] for n in names], allow_unused=True)))
    exp_grads = {k: (torch.zeros_like(ref[k]) if v is None else v) for k, v in exp_grads.items()}
    _assert_grads(got, exp_grads, torch.float64)
    if case == "mean":
        torch.testing.assert_close(got.uncertainty, torch.zeros_like(values["uncertainty"]))
    if case == "variance" and detach:
        for n in FIELDS[:-1]:
            torch.testing.assert_close(getattr(got, n), torch.zeros_like(values[n]))
        assert bool(torch.any(got.uncertainty != 0))


@pytest.mark.parametrize("op", ["rasterize", "voxelize"])
@pytest.mark.parametrize("use_beta,uncertainty", [(False, False), (True, True)])
def test_empty_primitives(device, op, use_beta, uncertainty):
    values = _scene(device, use_beta, uncertainty, requires_grad=True)
    for n in FIELDS:
        if values[n] is not None:
            values[n] = values[n].detach()[:0].requires_grad_()
    out, ctx = _forward(op, values)
    for v in _outputs(out):
        assert v.shape == ((2, 7, 9) if op == "rasterize" else (5, 7, 9))
        assert torch.count_nonzero(v) == 0
    got = getattr(r2splat_amd, op + "_bwd")(ctx, *_weights(out))
    _assert_grads(got, {n: torch.zeros(values[n].shape, dtype=torch.float64) for n in FIELDS
                        if values[n] is not None}, torch.float64)


@pytest.mark.parametrize("op", ["rasterize", "voxelize"])
def test_context_is_single_use(device, op):
    out, ctx = _forward(op, _scene(device))
    backward = getattr(r2splat_amd, op + "_bwd")
    backward(ctx, *_weights(out))
    with pytest.raises(RuntimeError):
        backward(ctx, *_weights(out))


@pytest.mark.parametrize("op,changed", [(op, n) for op in ("rasterize", "voxelize")
                                        for n in FIELDS + (("viewmats", "Ks") if op == "rasterize"
                                                           else ("volume_extent", "volume_center"))])
def test_context_rejects_in_place_changes(device, op, changed):
    values = _scene(device)
    backing = values["means"].new_empty(2, 6)
    backing[:, ::2] = values["means"]
    values["means"] = backing[:, ::2]
    out, ctx = _forward(op, values)
    values[changed].add_(0.01)
    with pytest.raises(RuntimeError):
        getattr(r2splat_amd, op + "_bwd")(ctx, *_weights(out))


@pytest.mark.parametrize("op", ["rasterize", "voxelize"])
@pytest.mark.parametrize("channel", [0, 1])
@pytest.mark.parametrize("bad", ["shape", "dtype", "device", "nonfinite"])
def test_bad_adjoints_do_not_consume_context(device, op, channel, bad):
    out, ctx = _forward(op, _scene(device))
    adj = _weights(out)
    v = adj[channel]
    if bad == "shape":
        inv = v.reshape(-1)
    elif bad == "dtype":
        inv = v.float()
    elif bad == "device":
        inv = v.cpu()
    else:
        inv = v.clone()
        inv.reshape(-1)[0] = float("nan")
    bad_adj = list(adj)
    bad_adj[channel] = inv
    backward = getattr(r2splat_amd, op + "_bwd")
    with pytest.raises((TypeError, ValueError)):
        backward(ctx, *bad_adj)
    assert bool(torch.isfinite(backward(ctx, *adj).densities).all())


@pytest.mark.parametrize("op,name", [("rasterize", "viewmats"), ("rasterize", "Ks"),
                                     ("voxelize", "volume_extent"), ("voxelize", "volume_center")])
def test_geometry_gradients_rejected(device, op, name):
    values = _scene(device)
    geom = values[name]
    storage = geom.new_empty((*geom.shape, 2))
    storage[..., 0] = geom
    values[name] = storage[..., 0].requires_grad_()
    with pytest.raises(ValueError, match=f"gradients with respect to {name}"):
        _forward(op, values)


@pytest.mark.parametrize("op", ["rasterize", "voxelize"])
def test_variance_adjoint_requires_uncertainty(device, op):
    out, ctx = _forward(op, _scene(device, uncertainty=False))
    w = _weights(out)[0]
    with pytest.raises(ValueError):
        getattr(r2splat_amd, op + "_bwd")(ctx, w, w)


def test_explicit_path_never_calls_autograd(device, monkeypatch):
    def forbidden(*a, **k):
        raise AssertionError("explicit path called autograd")

    values = _scene(device, requires_grad=True)
    monkeypatch.setattr(torch.autograd, "grad", forbidden)
    monkeypatch.setattr(torch.Tensor, "backward", forbidden)
    for op in ("rasterize", "voxelize"):
        out, ctx = _forward(op, values)
        got = getattr(r2splat_amd, op + "_bwd")(ctx, *_weights(out))
        for v in _outputs(out) + tuple(getattr(got, n) for n in FIELDS):
            assert not v.requires_grad and v.grad_fn is None and bool(torch.isfinite(v).all())
    assert all(values[n].grad is None for n in FIELDS)


@pytest.mark.parametrize("row", [(1e5, 0.002236059592288547, -4.999962500249999e-6),
                                 (1e7, 0.0002236067893647245, -4.999999625000025e-8)])
def test_large_float32_beta_centre_projection(device, row):
    gamma, ratio, dlog = row
    means = torch.tensor([[0., 0., 1.]], device=device)
    quats = means.new_tensor([[1., 0., 0., 0.]])
    scales = means.new_full((1, 3), 0.2)
    view = torch.eye(4, device=device)[None]
    K = means.new_tensor([[[1., 0., .5], [0., 1., .5], [0., 0., 1.]]])
    img, ctx = r2splat_amd.rasterize_fwd(means, quats, scales, means.new_tensor([1.]), view, K, (1, 1), eps=0,
                                         beta=means.new_tensor([gamma]))
    g = r2splat_amd.rasterize_bwd(ctx, torch.ones_like(img))
    expected = 0.2 * (2 * np.pi) ** 0.5 * ratio
    torch.testing.assert_close(img.cpu(), torch.full((1, 1, 1), expected), atol=0, rtol=4e-6)
    torch.testing.assert_close(g.densities.cpu(), torch.tensor([expected]), atol=0, rtol=4e-6)
    torch.testing.assert_close(g.beta.cpu(), torch.tensor([expected * dlog]), atol=0, rtol=4e-6)


# --------------------------------------------------------------------------- working-size parity
def _gpu_tile_mask(scene, eps, clip, tile=16):
    """Tile coverage of the operators in float64 (which the fused float32
    projection also computes in double)."""
    d = lambda t: t.double()  # noqa: E731
    covars, _ = ops.parameter_transform_fwd(d(scene.quats), d(scene.scales))
    m2, _, radii = ops.world_to_image_projection_fwd(d(scene.means), covars, d(scene.viewmats), d(scene.Ks),
                                                     scene.width, scene.height, eps, clip)
    return O.tile_mask_2d(m2.cpu().numpy(), radii.cpu().numpy(), tile, scene.width, scene.height)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_config0_rasterize_parity_with_tile_aware_oracle(device, dtype):
    """N = 1000, C = 4, 64 x 64, eps 1e-8, radius_clip 3: image within 1e-5 and
    every parameter gradient within 1e-4 (float32) of the float64 oracle
    restricted to the intersected tiles, as max|a-b| / max|b|."""
    sc = workloads.config0(device=device, dtype=dtype)
    mask = _gpu_tile_mask(sc, 1e-8, 3.0)
    img, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size)
    leaves = [t.detach().cpu().double().requires_grad_() for t in sc.params()]
    ref = O.rasterize_scene(*leaves, sc.viewmats.cpu().double(), sc.Ks.cpu().double(), sc.image_size, eps=1e-8,
                            mask=mask)
    t_img = 1e-5 if dtype == torch.float32 else 1e-10
    err = O.relative_error(img, ref)
    assert err <= t_img, f"image relative error {err:.2e}"
    v = (2 * (img - 0.8 * img) / img.numel()).contiguous()
    grads = r2splat_amd.rasterize_bwd(ctx, v)
    exp = torch.autograd.grad((ref * v.cpu().double()).sum(), leaves)
    t_grad = 1e-4 if dtype == torch.float32 else 1e-8
    for name, e in zip(("means", "quats", "scales", "densities"), exp):
        err = O.relative_error(getattr(grads, name), e)
        assert err <= t_grad, f"{name} relative error {err:.2e}"


def test_config0_voxelize_parity(device):
    sc = workloads.config0(device=device, dtype=torch.float32)
    shape = (24, 24, 24)
    vol, ctx = r2splat_amd.voxelize_fwd(*sc.params(), shape, volume_extent=(1.2, 1.2, 1.2))
    covars, mx = ops.parameter_transform_fwd(sc.quats, sc.scales)
    ext = torch.tensor([1.2, 1.2, 1.2], device=device)
    mapped, _, radii = ops.world_to_volume_mapping_fwd(sc.means, covars, mx, ext, torch.zeros(3, device=device),
                                                       24, 24, 24, 1e-8, 3.0)
    mask = O.tile_mask_3d(mapped.cpu().numpy(), radii.cpu().numpy(), 8, shape)
    leaves = [t.detach().cpu().double().requires_grad_() for t in sc.params()]
    ref = O.voxelize(*leaves, shape, (1.2, 1.2, 1.2), (0, 0, 0), eps=1e-8, mask=mask)
    assert O.relative_error(vol, ref) <= 1e-5
    v = torch.linspace(-1, 1, vol.numel(), device=device).reshape(shape)
    grads = r2splat_amd.voxelize_bwd(ctx, v)
    exp = torch.autograd.grad((ref * v.cpu().double()).sum(), leaves)
    for name, e in zip(("means", "quats", "scales", "densities"), exp):
        assert O.relative_error(getattr(grads, name), e) <= 1e-4, name

"""Graph-captured fits (r2splat_amd.train.GaussianFit): one step equals the
explicit path bitwise, graph replays equal eager steps, Adam matches
torch.optim.Adam, the MSE value matches torch, and a capacity overflow inside
captured steps is skipped on the device and replayed so the final parameters
equal those of a run with ample capacity. Also: autograd-wrapper steps
captured with torch.cuda.graph equal eager ones."""

import pytest
import torch

import r2splat_amd
from r2splat_amd import autograd as ag
from r2splat_amd import workloads
from r2splat_amd.train import AdamConfig, GaussianFit

pytestmark = pytest.mark.gpu


def _scene(device, n=1500, c=4, h=64, w=80, seed=0):
    sc = workloads.scene(n, c, h, w, scale_range=(0.01, 0.08), seed=seed, device=device)
    tgt = workloads.scene(n, c, h, w, scale_range=(0.01, 0.08), seed=seed + 1, device=device)
    target, _ = r2splat_amd.rasterize_fwd(*tgt.params(), tgt.viewmats, tgt.Ks, tgt.image_size)
    return sc, target


def _explicit_grads(sc, target, params=None):
    params = params or sc.params()
    img, ctx = r2splat_amd.rasterize_fwd(*params, sc.viewmats, sc.Ks, sc.image_size)
    v = (2.0 / img.numel()) * (img - target)
    return img, r2splat_amd.rasterize_bwd(ctx, v)


def test_fit_step_equals_explicit_path(device):
    sc, target = _scene(device)
    fit = GaussianFit(*sc.params(), sc.viewmats, sc.Ks, target)
    fit.step(1, graph=False)
    fit.sync()
    img, ref = _explicit_grads(sc, target)
    assert torch.equal(fit.image, img)
    for name, got in zip(("means", "quats", "scales", "densities"),
                         (fit.v_means, fit.v_quats, fit.v_scales, fit.v_densities)):
        assert torch.equal(got, getattr(ref, name)), name
    expect = float(((img.double() - target.double()) ** 2).mean())
    assert abs(fit.loss_value() - expect) <= 1e-6 * expect


def test_graph_replays_equal_eager_steps(device):
    sc, target = _scene(device, seed=3)
    fit = GaussianFit(*sc.params(), sc.viewmats, sc.Ks, target)
    fit.step(1, graph=False)
    fit.sync()
    eager = [t.clone() for t in (fit.image, fit.grads, fit.loss)]
    fit.capture()
    for k in range(4):
        for t in (fit.image, fit.grads, fit.loss):
            t.fill_(float("nan"))
        fit.step(1)
        fit.sync()
        # the program's own eager work between replays
        _explicit_grads(sc, target)
        torch.cuda.synchronize()
        for a, b in zip((fit.image, fit.grads, fit.loss), eager):
            assert torch.equal(a, b), f"replay {k}"


def test_adam_matches_torch_optimizer(device):
    sc, target = _scene(device, n=800, seed=5)
    sc.densities.mul_(0.8).add_(0.2)  # keeps the explicit path's densities >= 0 check satisfied under updates
    cfg = AdamConfig(lr=1e-3)
    fit = GaussianFit(*sc.params(), sc.viewmats, sc.Ks, target, adam=cfg)
    fit.capture()
    fit.step(5)
    assert fit.sync() == 5
    params = [t.detach().clone() for t in sc.params()]
    opt = torch.optim.Adam(params, lr=cfg.lr, betas=(cfg.beta1, cfg.beta2), eps=cfg.eps, foreach=False)
    for _ in range(5):
        _, g = _explicit_grads(sc, target, params)
        for p, gp in zip(params, (g.means, g.quats, g.scales, g.densities)):
            p.grad = gp
        opt.step()
    for got, exp in zip((fit.means, fit.quats, fit.scales, fit.densities), params):
        err = float((got - exp).abs().max() / exp.abs().max())
        assert err <= 2e-6, err


def test_capacity_overflow_is_skipped_and_replayed(device):
    sc, target = _scene(device, n=1200, seed=7)
    cfg = AdamConfig(lr=1e-3)
    ample = GaussianFit(*sc.params(), sc.viewmats, sc.Ks, target, adam=cfg)
    ample.capture()
    ample.step(3)
    assert ample.sync() == 3
    small = GaussianFit(*sc.params(), sc.viewmats, sc.Ks, target, adam=cfg, capacity=1024)
    small.capture()
    small.step(3)
    assert int(small.adam_state[1].item()) == 3  # every step overflowed and was skipped
    assert small.sync() == 3
    assert small.capacity > 1024
    assert torch.equal(small.params, ample.params)
    assert torch.equal(small.exp_avg, ample.exp_avg)


def test_captured_autograd_wrapper_step_equals_eager(device):
    sc, target = _scene(device, seed=9)
    params = [t.detach().clone().requires_grad_() for t in sc.params()]

    def step():
        img = ag.rasterize(*params, sc.viewmats, sc.Ks, sc.image_size, recompute=False)
        return torch.autograd.grad(((img - target) ** 2).mean(), params)

    eager = [g.clone() for g in step()]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = step()
    for k in range(3):
        graph.replay()
        torch.cuda.synchronize()
        step()  # eager work between replays
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            assert torch.equal(a, b), f"replay {k}"


def test_beta_fit_matches_explicit_path_and_torch_adam(device):
    """Beta primitives: one captured step equals the explicit API bitwise
    (image and all five gradients), and captured Adam steps follow
    torch.optim.Adam on the explicit path's gradients."""
    sc = workloads.scene(900, 3, 56, 64, scale_range=(0.02, 0.1), seed=13, device=device, beta_range=(0.5, 4.0))
    tg = workloads.scene(900, 3, 56, 64, scale_range=(0.02, 0.1), seed=14, device=device, beta_range=(0.5, 4.0))
    target, _ = r2splat_amd.rasterize_fwd(*tg.params(), tg.viewmats, tg.Ks, tg.image_size, beta=tg.beta)
    fit = GaussianFit(*sc.params(), sc.viewmats, sc.Ks, target, beta=sc.beta)
    fit.capture()
    fit.step(1)
    fit.sync()
    img, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size, beta=sc.beta)
    ref = r2splat_amd.rasterize_bwd(ctx, (2.0 / img.numel()) * (img - target))
    assert torch.equal(fit.image, img)
    for name, got in zip(("means", "quats", "scales", "densities", "beta"),
                         (fit.v_means, fit.v_quats, fit.v_scales, fit.v_densities, fit.v_beta)):
        assert torch.equal(got, getattr(ref, name)), name

    cfg = AdamConfig(lr=1e-3, min_scale=1e-6)
    fit = GaussianFit(*sc.params(), sc.viewmats, sc.Ks, target, beta=sc.beta, adam=cfg)
    fit.capture()
    fit.step(4)
    assert fit.sync() == 4
    params = [t.detach().clone() for t in sc.params() + (sc.beta,)]
    opt = torch.optim.Adam(params, lr=cfg.lr, betas=(cfg.beta1, cfg.beta2), eps=cfg.eps, foreach=False)
    for _ in range(4):
        img, ctx = r2splat_amd.rasterize_fwd(*params[:4], sc.viewmats, sc.Ks, sc.image_size, beta=params[4])
        g = r2splat_amd.rasterize_bwd(ctx, (2.0 / img.numel()) * (img - target))
        for p, gp in zip(params, (g.means, g.quats, g.scales, g.densities, g.beta)):
            p.grad = gp
        opt.step()
        with torch.no_grad():
            params[2].clamp_(min=1e-6)
            params[3].clamp_(min=0.0)
            params[4].clamp_(min=cfg.min_beta)
    for got, exp in zip((fit.means, fit.quats, fit.scales, fit.densities, fit.beta), params):
        err = float((got - exp).abs().max() / exp.abs().max())
        assert err <= 2e-6, err

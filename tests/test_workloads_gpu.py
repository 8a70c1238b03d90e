"""The batched and serving workloads at their real job mix.

* Batched (32 uneven jobs: N log-uniform in [10k, 400k], C in {16, 32, 64,
  128} views, side in {128, 256, 512}; fused fwd + MSE + bwd + Adam lr 1e-2,
  one hipGraph per batched step, tile size and waves per job from the
  autotuner): after a few steps every job's parameters equal, bit for bit,
  the same job fitted alone with the same kernel configuration. The batched
  step time is reported against the same jobs stepped one by one on the plain
  operator path.
* Serving (Poisson arrivals at 4 requests/s of the same job mix, Gaussian or
  Beta with p = 0.5, over a shortened window and fewer iterations per
  request): every request finishes, p50 / p99 time-to-finished-fit are
  finite, and a sample of requests ends bitwise equal to the request fitted
  alone.
"""

import math
import time

import pytest
import torch

from r2splat_amd import autograd as ag
from r2splat_amd import explicit, workloads
from r2splat_amd.batched import Autotuner, BatchedFit
from r2splat_amd.serving import FitServer, build_job, job_config, poisson_requests
from r2splat_amd.train import AdamConfig, GaussianFit

pytestmark = pytest.mark.gpu

ADAM = AdamConfig(lr=1e-2, min_scale=1e-6)
STEPS = 3


def _alone(job, adam, steps, **config):
    fit = GaussianFit(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"],
                      job["target"], adam=adam, beta=job.get("beta"), **config)
    fit.capture()
    fit.step(steps)
    fit.sync()
    out = fit.params.clone()
    del fit
    return out


def _plain_step_ms(job):
    """One step of the job through the autograd wrapper on the staged operator path with torch Adam."""
    params = [job[k].detach().clone().requires_grad_() for k in ("means", "quats", "scales", "densities")]
    opt = torch.optim.Adam(params, lr=ADAM.lr)
    old = explicit.BACKEND
    explicit.BACKEND = "ops"
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        img = ag.rasterize(*params, job["viewmats"], job["Ks"], tuple(job["target"].shape[-2:]))
        loss = ((img - job["target"]) ** 2).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3
    finally:
        explicit.BACKEND = old


def test_config3_batched_jobs_each_equal_the_job_fitted_alone(device):
    specs = workloads.config3_specs(32, seed=0)
    print(f"batched: {len(specs)} jobs {[s[:3] for s in specs]}", flush=True)
    jobs = [workloads.fit_job(n, c, side, seed, device=device) for n, c, side, seed in specs]
    batch = BatchedFit(jobs, adam=ADAM, n_streams=4, tuner=Autotuner(repeats=1))
    batch.step(1)
    batch.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.step(STEPS - 1)
    assert batch.sync() == STEPS
    batched_ms = (time.perf_counter() - t0) * 1e3 / (STEPS - 1)
    for j, (job, fit) in enumerate(zip(jobs, batch.fits)):
        config = dict(tile_size=fit.tile_size, fwd_waves=fit.fwd_waves, bwd_waves=fit.bwd_waves)
        assert torch.equal(fit.params, _alone(job, ADAM, STEPS, **config)), (j, specs[j][:3], config)
    plain_ms = sum(_plain_step_ms(job) for job in jobs)
    print(f"batched step {batched_ms:.1f} ms (32 jobs, one graph) vs plain path {plain_ms:.1f} ms: "
          f"{plain_ms / batched_ms:.2f}x", flush=True)


def test_config4_serving_window_latency_and_alone_vs_mix(device):
    reqs = poisson_requests(4.0, 4.0, seed=0, steps=20)
    print(f"serving: {len(reqs)} requests {[(r.n, r.views, r.side, r.beta) for r in reqs]}", flush=True)
    server = FitServer(device, n_streams=4, adam=ADAM)
    stats = server.run(reqs, keep_results=True)
    assert sorted(f[0] for f in stats.finished) == [r.rid for r in reqs]
    p50, p99 = stats.percentile(50), stats.percentile(99)
    print(f"serving: p50 {p50:.2f} s, p99 {p99:.2f} s, {stats.job_steps / stats.wall:.0f} served iters/s, "
          f"peak {stats.peak_live} live jobs", flush=True)
    assert math.isfinite(p50) and math.isfinite(p99) and 0 < p50 <= p99
    del server
    sample = reqs[:2] + [r for r in reqs if r.beta][:1]
    for req in sample:
        cfg = job_config(req)
        config = dict(tile_size=cfg.tile_size, fwd_waves=cfg.fwd_waves, bwd_waves=cfg.bwd_waves)
        assert torch.equal(stats.results[req.rid], _alone(build_job(req, device), ADAM, req.steps, **config)), req.rid

"""Batched graph-captured fits (r2splat_amd.batched): every job's result in the
batch equals the job fitted alone (bitwise), whatever stream and shared
workspace it landed on; per-job kernel configurations; the autotuner; a job
overflowing its capacity inside the batch is repaired without disturbing
the others."""

import pytest
import torch

from r2splat_amd import workloads
from r2splat_amd.batched import Autotuner, BatchedFit, KernelConfig
from r2splat_amd.train import AdamConfig, GaussianFit

pytestmark = pytest.mark.gpu

SPECS = [(900, 3, 48, 11), (400, 2, 64, 13), (1500, 4, 40, 15), (250, 5, 56, 17)]
CONFIGS = [KernelConfig(16, 8, 8), KernelConfig(8, 4, 16), KernelConfig(32, 16, 4), KernelConfig(16, 4, 4)]


def _jobs(device, capacities=None):
    jobs = []
    for k, (n, c, side, seed) in enumerate(SPECS):
        job = workloads.fit_job(n, c, side, seed, device=device, scale_range=(0.01, 0.08))
        job["config"] = CONFIGS[k]
        if capacities:
            job["capacity"] = capacities[k]
        jobs.append(job)
    return jobs


def _alone(job, adam, steps):
    cfg = job["config"]
    fit = GaussianFit(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"],
                      job["target"], adam=adam, tile_size=cfg.tile_size, fwd_waves=cfg.fwd_waves,
                      bwd_waves=cfg.bwd_waves)
    fit.capture()
    fit.step(steps)
    assert fit.sync() == steps
    return fit


def test_batched_jobs_equal_jobs_fitted_alone(device):
    adam = AdamConfig(lr=1e-3)
    jobs = _jobs(device)
    batch = BatchedFit(jobs, adam=adam, n_streams=2)
    batch.step(3)
    assert batch.sync() == 3
    for job, fit in zip(jobs, batch.fits):
        ref = _alone(job, adam, 3)
        assert torch.equal(fit.params, ref.params)
        assert torch.equal(fit.exp_avg_sq, ref.exp_avg_sq)
        assert torch.equal(fit.loss, ref.loss)


def test_batched_overflow_in_one_job_is_repaired(device):
    adam = AdamConfig(lr=1e-3)
    jobs = _jobs(device, capacities=[None, 1024, None, None])
    batch = BatchedFit(jobs, adam=adam, n_streams=2)
    batch.step(2)
    assert batch.sync() == 2
    assert batch.fits[1].capacity > 1024
    for job, fit in zip(_jobs(device), batch.fits):
        ref = _alone(job, adam, 2)
        assert torch.equal(fit.params, ref.params)


def test_autotuner_times_every_candidate_and_caches(device):
    job = workloads.fit_job(3000, 4, 96, 21, device=device)
    tuner = Autotuner(repeats=1)
    rec = tuner.tune(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"], 96, 96)
    assert rec.config.tile_size in (8, 16, 32) and rec.config.fwd_waves in (4, 8, 16)
    assert len(rec.timings) == 3 * (1 + 2 * 3)
    best = min(rec.timings[(t, "prepare", 0)] + min(rec.timings[(t, "render", w)] for w in (4, 8, 16))
               + min(rec.timings[(t, "raster_vjp", w)] for w in (4, 8, 16)) for t in (8, 16, 32))
    assert abs(rec.ms - best) <= 1e-9
    assert tuner.tune(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"], 96,
                      96) is rec
    # a tuned batch runs and matches its jobs alone
    jobs = [dict(job)]
    batch = BatchedFit(jobs, adam=AdamConfig(lr=1e-3), n_streams=1, tuner=tuner)
    assert batch.fits[0].tile_size == rec.config.tile_size
    batch.step(1)
    assert batch.sync() == 1


def test_mixed_gaussian_and_beta_batch_equals_jobs_alone(device):
    adam = AdamConfig(lr=1e-3, min_scale=1e-6)
    jobs = []
    for k, (n, c, side, seed) in enumerate(SPECS[:3]):
        job = workloads.fit_job(n, c, side, seed, device=device, scale_range=(0.01, 0.08),
                                beta_range=(0.5, 4.0) if k != 1 else None)
        job["config"] = CONFIGS[k]
        jobs.append(job)
    batch = BatchedFit(jobs, adam=adam, n_streams=2)
    batch.step(3)
    assert batch.sync() == 3
    for job, fit in zip(jobs, batch.fits):
        cfg = job["config"]
        ref = GaussianFit(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"],
                          job["target"], adam=adam, tile_size=cfg.tile_size, fwd_waves=cfg.fwd_waves,
                          bwd_waves=cfg.bwd_waves, beta=job["beta"])
        ref.capture()
        ref.step(3)
        assert ref.sync() == 3
        assert torch.equal(fit.params, ref.params)

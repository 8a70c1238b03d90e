"""Batched graph-captured fits (r2splat_amd.batched): every job's result in the
batch equals the job fitted alone (bitwise), with the raster stages packed
into shared launches or each job's step on its stream (sharing a workspace);
per-job kernel configurations; the autotuner; a job overflowing its capacity
inside the batch is repaired without disturbing the others; packed raster
launches equal each job's own launches bitwise."""

import ctypes

import pytest
import torch

from r2splat_amd import fused, workloads
from r2splat_amd.batched import Autotuner, BatchedFit, KernelConfig
from r2splat_amd.hip import _backend as be
from r2splat_amd.hip._backend import call, ptr
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


@pytest.mark.parametrize("packed", [True, False])
def test_batched_jobs_equal_jobs_fitted_alone(device, packed):
    adam = AdamConfig(lr=1e-3)
    jobs = _jobs(device)
    batch = BatchedFit(jobs, adam=adam, n_streams=2, packed=packed)
    batch.step(3)
    assert batch.sync() == 3
    for job, fit in zip(jobs, batch.fits):
        ref = _alone(job, adam, 3)
        assert torch.equal(fit.params, ref.params)
        assert torch.equal(fit.exp_avg_sq, ref.exp_avg_sq)
        assert torch.equal(fit.loss, ref.loss)


@pytest.mark.parametrize("packed", [True, False])
def test_batched_overflow_in_one_job_is_repaired(device, packed):
    adam = AdamConfig(lr=1e-3)
    jobs = _jobs(device, capacities=[None, 1024, None, None])
    batch = BatchedFit(jobs, adam=adam, n_streams=2, packed=packed)
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


@pytest.mark.parametrize("use_beta,use_var,detach", [(False, False, 0), (True, False, 0), (False, True, 0),
                                                     (True, True, 1)])
def test_packed_raster_launches_equal_each_jobs_own(device, use_beta, use_var, detach):
    """Four jobs of different views, sizes and image shapes (one of them
    overflowed, rendering nothing) in one packed render and one packed raster
    VJP, for each record kind (with uncertainty: the variance image, its
    adjoint and detach): every image and every intersection partial equals
    the job's own launch bitwise."""
    specs = [(900, 3, 48, 40), (400, 2, 64, 64), (1500, 4, 40, 72), (250, 5, 56, 24)]
    T, waves = 16, 8
    stream = be.stream_of(device)
    preps, table = [], (be.RasterJob * len(specs))()
    first = 0
    for k, (n, c, h, w) in enumerate(specs):
        sc = workloads.scene(n, c, h, w, scale_range=(0.01, 0.08), seed=30 + k, device=device,
                             beta_range=(0.5, 4.0) if use_beta else None, uncertainty_max=0.1 if use_var else None)
        f = fused.prepare(*sc.params(), sc.viewmats, sc.Ks, w, h, 1e-8, 1.0 if use_beta else 3.0, tile_size=T,
                          beta=sc.beta, uncertainty=sc.uncertainty, capacity=64 if k == 2 else None, check=k != 2)
        kind = f.kind
        g = torch.Generator(device=device).manual_seed(k)
        v = torch.randn((c, h, w), generator=g, device=device)
        vv = torch.randn((c, h, w), generator=g, device=device) if use_var else None
        nan = lambda shape: torch.full(shape, float("nan"), device=device)  # noqa: E731
        mine = (nan((c, h, w)), nan((f.capacity * 8,)), nan((c, h, w)) if use_var else None)
        packed_out = (nan((c, h, w)), nan((f.capacity * 8,)), nan((c, h, w)) if use_var else None)
        call("r2s_fused_render", ptr(f.packed), ptr(f.offsets), ptr(f.tile_order), ptr(f.counters), f.capacity, c,
             w, h, T, waves, kind, ptr(mine[0]), ptr(mine[2]), stream)
        call("r2s_fused_raster_vjp", ptr(f.packed), ptr(f.offsets), ptr(f.tile_order), ptr(f.counters),
             f.capacity, c, w, h, T, waves, kind, ptr(v), ptr(vv), detach, ptr(mine[1]), stream)
        e = table[k]
        e.packed, e.tile_offsets, e.tile_order, e.counters = ptr(f.packed), ptr(f.offsets), ptr(f.tile_order), \
            ptr(f.counters)
        e.image, e.image_var, e.v_image, e.v_image_var = ptr(packed_out[0]), ptr(packed_out[2]), ptr(v), ptr(vv)
        e.isect_grads, e.detach = ptr(packed_out[1]), detach
        e.capacity, e.C, e.width, e.height, e.first_tile = f.capacity, c, w, h, first
        first += c * (-(-h // T)) * (-(-w // T))
        preps.append((f, mine, packed_out, (v, vv)))  # the table points into every job's buffers: keep them
    assert int(preps[2][0].counters[1].item()) == 1  # the overflowed job
    dev_table = torch.frombuffer(bytearray(ctypes.string_at(ctypes.addressof(table), ctypes.sizeof(table))),
                                 dtype=torch.uint8).to(device)
    call("r2s_fused_render_packed", ctypes.addressof(table), ptr(dev_table), len(specs), T, waves, kind, stream)
    call("r2s_fused_raster_vjp_packed", ctypes.addressof(table), ptr(dev_table), len(specs), T, waves, kind, stream)
    torch.cuda.synchronize(device)
    bits = lambda t: t.view(torch.int32)  # noqa: E731 (bitwise, NaN fill included)
    for k, (f, mine, packed_out, _) in enumerate(preps):
        assert torch.equal(bits(packed_out[0]), bits(mine[0])), k
        if use_var:
            assert torch.equal(bits(packed_out[2]), bits(mine[2])), k
        # slots of footprint-rectangle tiles the ellipse does not reach are never
        # written: both launches leave the same NaN fill there
        n = int(f.counters[0].item()) if k != 2 else 0
        assert torch.equal(bits(packed_out[1][: 8 * n]), bits(mine[1][: 8 * n])), k
    assert not bool(preps[2][2][0].any())  # an overflowed job renders zeros
    if use_var:
        assert not bool(preps[2][2][2].any())
    # a table whose first_tile prefix is wrong is refused before any launch
    table[1].first_tile += 1
    with pytest.raises(RuntimeError, match="starts at tile"):
        call("r2s_fused_render_packed", ctypes.addressof(table), ptr(dev_table), len(specs), T, waves, kind, stream)

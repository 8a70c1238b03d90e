"""Benchmark: fwd+bwd iterations/s of the CT splatting path on one MI355X.

Workload (one step): N = 100,000 Gaussians, 64 views of 256 x 256 (tile 16),
float32, through the torch.autograd.Function wrapper: rasterize forward -> MSE
against a target rendered from a second Gaussian set -> backward to the
physical parameter gradients. Inputs are resident in HBM before timing starts.
Voxelization of the same Gaussians to 128^3 (extent 1, tile 8) is timed
separately (fwd + bwd of a fixed random adjoint).

Prints one JSON line (metric, value, unit, timing, workload config), including the
roofline of the dominant kernel measured with HIP events on its launch stream
during the timed region, and a CPU baseline (the C/OpenMP restatement of the
same iteration timed on a sample of the views on the host cores).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+bwd iters/s (100k Gauss, 64×256² views) vs plain path; p99 fit latency"
HBM_PEAK_GBS = 8000.0
# FP32 vector peak without packed FP32 (which every kernel here is built
# without): 256 CUs x 4 SIMDs x 16 lanes x 2 FLOP per FMA x 2.4 GHz.
FP32_PEAK_TFLOPS = 78.6
# Reference-equivalent FP32 operations per pixel-primitive pair that passes the
# cut-off, counted from the reference's per-pair arithmetic with an exp as one
# operation: forward 16 (delta 2, sigma 9, negate+exp 2, opacity*mu*e 2,
# accumulate 1; RasterizeToPixelsFwd.cu:160-179), backward 46 (sigma
# recompute 13, alpha 2, v_sigma 2, v_xy 12, v_conic 11, v_opacity 3, v_mu 3;
# RasterizeToPixelsBwd.cu:430-468). Pairs the cut-off drops contribute nothing
# and are not counted as work.
FLOP_PER_PAIR = {"rasterize_fwd": 16.0, "rasterize_bwd": 46.0}
# serving: the plain per-job path is compared with the server's step over this many iterations
PLAIN_CHECK_STEPS = 5
# Vector (VALU) instructions one launch of each fused raster kernel issues on
# this workload (PMC SQ_INSTS_VALU, summed over the device, from
# `rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU ... --kernel-include-regex fused_raster`
# on `bench.py --steps 2 --warmup 1 --plain-steps 0 --no-cpu-baseline --no-voxel
# --no-graph`; profiles/raster_100k_64x256_occupancy_valu.md). The kernels
# evaluate only the rows (backward) and row groups (forward) each record can
# reach, so they issue far fewer operations than pairs x FLOP_PER_PAIR.
VALU_INSTS = {"rasterize_fwd": 7.333e8, "rasterize_bwd": 8.121e8}
SIMDS = 256 * 4          # compute units x SIMDs
CLOCK_HZ = 2.4e9         # peak engine clock
# Bytes each launch of the raster kernels reads and writes beyond L2 on this
# workload, from the L2-to-memory requests counted by size in rocprofv3 --pmc
# runs of `bench.py --steps 2 --warmup 1 --plain-steps 0 --no-cpu-baseline
# --no-voxel --no-graph --serve-window 0 --no-parity --copy-calibration`:
# reads 128 TCC_EA0_RDREQ_128B + 64 TCC_EA0_RDREQ_64B + 32 TCC_EA0_RDREQ_32B,
# writes 64 TCC_EA0_WRREQ_64B + 32 (TCC_EA0_WRREQ - TCC_EA0_WRREQ_64B)
# (medians over the dispatches). The plain 256 MiB copy of the same run
# counts exactly its 268,435,456 bytes each way this way; FETCH_SIZE counts
# its 128-byte reads as 64 bytes (half). Summary and commands:
# profiles/raster_100k_64x256_hbm_traffic.md.
TRAFFIC_BYTES = {"rasterize_fwd": (2_303_045_248, 17_313_472), "rasterize_bwd": (2_398_399_040, 921_594_144)}


def stage_rooflines(stages_ms, N, CN, I, kept_fwd, kept_bwd):
    """Per-stage roofline of the step against the compulsory work of the
    reference's stage decomposition (FP32): HBM bytes for the stages that
    stream memory, FLOP on the pairs that pass the cut-off for the raster
    stages; and the iteration floor (the stages at their peaks, one after the
    other) against the measured stage sum. Bytes per stage: projection 36N +
    40CN, conics 40CN, tile count 20CN, scan 12CN, generate 24CN + 12I, sort
    2 x 32I, offsets 8I (binning: 36N + 136CN + 84I); conics VJP 80CN,
    projection VJP 32CN + 72N, parameter VJP 80N, density reduce 4CN
    (parameter VJP: 152N + 116CN)."""
    work = {"project_bin": ("hbm", 36 * N + 136 * CN + 84 * I),
            "rasterize_fwd": ("valu", kept_fwd * FLOP_PER_PAIR["rasterize_fwd"]),
            "rasterize_bwd": ("valu", kept_bwd * FLOP_PER_PAIR["rasterize_bwd"]),
            "param_bwd": ("hbm", 152 * N + 116 * CN)}
    out, floor_ms, total_ms = {}, 0.0, 0.0
    for name, (bound, amount) in work.items():
        if name not in stages_ms:
            continue
        ms = stages_ms[name][0]
        if bound == "hbm":
            achieved, peak, unit = amount / (ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s"
            floor = amount / (HBM_PEAK_GBS * 1e9) * 1e3
        else:
            achieved, peak, unit = amount / (ms * 1e-3) / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s"
            floor = amount / (FP32_PEAK_TFLOPS * 1e12) * 1e3
        out[name] = {"bound": bound, "achieved": round(achieved, 1), "peak": peak, "unit": unit,
                     "frac": round(achieved / peak, 3), "ms": round(ms, 4)}
        floor_ms += floor
        total_ms += ms
    out["iteration"] = {"floor_ms": round(floor_ms, 3), "stage_ms": round(total_ms, 3),
                        "frac": round(floor_ms / total_ms, 3) if total_ms else None}
    return out


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--path", default="explicit", choices=["explicit"])
    p.add_argument("--plain-steps", type=int, default=5, help="timed steps of the staged operator path")
    p.add_argument("--cpu-views", type=int, default=64, help="views in the CPU-baseline sample")
    p.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU-baseline sample time")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-voxel", action="store_true", help="skip the separate 128^3 voxelization timing")
    p.add_argument("--no-graph", action="store_true", help="time the eager step instead of its hipGraph replay")
    p.add_argument("--no-parity", action="store_true", help="skip the float64 error report of the timed workload")
    p.add_argument("--serve-window", type=float, default=10.0,
                   help="config1: seconds of Poisson fit requests (4 req/s) served for the p50/p99 fit latency "
                        "(0 skips it)")
    p.add_argument("--workload", default="config1", choices=["config1", "config2", "batched", "serving"],
                   help="config1: the metric workload; batched: 32 uneven jobs with Adam in one graph; "
                        "serving: Poisson-arriving fit requests through the continuous-batching server; "
                        "config2: Beta primitives with uncertainty (mean, variance) and a 256^3 volume with TV")
    p.add_argument("--rate", type=float, default=4.0, help="serving: requests per second")
    p.add_argument("--duration", type=float, default=120.0, help="serving: arrival window in seconds")
    p.add_argument("--serve-steps", type=int, default=200, help="serving: Adam iterations per request")
    p.add_argument("--plain-jobs", type=int, default=3, help="serving: requests re-fitted on the plain path")
    p.add_argument("--jobs", type=int, default=32, help="jobs of the batched workload")
    p.add_argument("--streams", type=int, default=4, help="streams of the batched workload")
    p.add_argument("--max-active", type=int, default=0,
                   help="serving: step only the earliest-arrived N live jobs per tick (0: every live job)")
    p.add_argument("--active-ms", type=float, default=0.0,
                   help="serving: step the earliest-arrived live jobs whose estimated iteration times add up to "
                        "this many ms per tick (0: every live job)")
    p.add_argument("--no-packed", action="store_true",
                   help="batched: run each job's raster stages as its own launches on its stream instead of "
                        "packed launches over all jobs' tiles")
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run one plain 256 MiB device copy (reference kernel for HBM byte counters)")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this implementation runs on one GPU; --gpus must be 1 (sharding across "
                                   "GPUs is a design, not a measured path)"}))
        sys.exit(2)
    import torch

    import r2splat_amd
    from r2splat_amd import profiling, workloads

    if args.workload == "batched":
        return batched_main(args)
    if args.workload == "serving":
        return serving_main(args)
    if args.workload == "config2":
        return config2_main(args)
    device = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones((64 << 20,), dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    sc = workloads.config1(device=device, dtype=torch.float32, seed=0)
    tgt_scene = workloads.config1(device=device, dtype=torch.float32, seed=1)
    with torch.no_grad():
        target, _ = r2splat_amd.rasterize_fwd(*tgt_scene.params(), tgt_scene.viewmats, tgt_scene.Ks, sc.image_size)
    del tgt_scene

    from r2splat_amd import autograd as ag

    params = [t.detach().requires_grad_() for t in sc.params()]

    def step():
        img = ag.rasterize(*params, sc.viewmats, sc.Ks, sc.image_size, recompute=False)
        loss = ag.mse_loss(img, target)
        return torch.autograd.grad(loss, params)

    def timed(backend, steps, warmup):
        """(ms per step without stage events, stage times from a second pass
        of the same steps with them)."""
        explicit.BACKEND = backend
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        profiling.reset()
        profiling.enable(True)
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        profiling.enable(False)
        return (t1 - t0) * 1e3 / steps, profiling.collect()

    from r2splat_amd import explicit

    plain = timed("ops", args.plain_steps, 1) if args.plain_steps > 0 else None
    eager_ms, stages = timed("fused", args.steps, args.warmup)
    ms = graph_timing(step, args.steps, args.warmup) if not args.no_graph else eager_ms

    # intersections of this workload (fixed across steps: the parameters do not change)
    from r2splat_amd import hip as ops

    pc = pair_counts(sc)
    n_isects = pc["intersections"]
    kernel_stages = {k: v for k, v in stages.items() if k in FLOP_PER_PAIR}
    dominant = max(kernel_stages, key=lambda k: kernel_stages[k][0])
    dom_ms = kernel_stages[dominant][0]
    kept = pc["kept_bwd"] if dominant == "rasterize_bwd" else pc["kept_fwd"]
    achieved = kept * FLOP_PER_PAIR[dominant] / (dom_ms * 1e-3) / 1e12
    issue = VALU_INSTS[dominant] / (dom_ms * 1e-3 * CLOCK_HZ * SIMDS)
    traffic = sum(TRAFFIC_BYTES[dominant])
    roofline = {"bound": "valu", "kernel": dominant, "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "traffic_gbs": round(traffic / (dom_ms * 1e-3) / 1e9, 1),
                "algorithmic_work": f"{kept:.4e} pixel-primitive pairs that pass the cut-off x "
                                    f"{FLOP_PER_PAIR[dominant]:.0f} FLOP (of {pc['list_pairs']:.4e} pairs in the "
                                    f"intersected tiles)",
                "avg_launch_ms": round(dom_ms, 4),
                "stages": stage_rooflines(stages, sc.means.shape[0], sc.viewmats.shape[0] * sc.means.shape[0],
                                          n_isects, pc["kept_fwd"], pc["kept_bwd"]),
                "issued": {"valu_insts": VALU_INSTS[dominant], "per_simd_cycle": round(issue, 4),
                           "note": "vector instructions the kernel issues per SIMD per cycle (measured count / "
                                   "(launch time x 2.4 GHz x 1,024 SIMDs)); a 16-lane SIMD completes one "
                                   "full-rate wave64 instruction per 4 cycles"}}

    parity = None if args.no_parity else parity_vs_fp64(sc, target)
    voxel = None if args.no_voxel else voxel_timing(sc, args.steps)
    latency = fit_latency(args.rate, args.serve_window, args.serve_steps, args.streams, args.max_active or None,
                          args.active_ms or None) if args.serve_window > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, target, args.cpu_views, args.cpu_seconds)

    result = {
        "metric": METRIC,
        "value": round(1e3 / ms, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian sets, orbit cameras; target from a second seeded set)",
        "config": {"workload": "config1: N=100000 Gaussians, C=64 views of 256x256, tile 16, MSE fwd+bwd "
                               "through rasterize_fwd/rasterize_bwd",
                   "path": args.path, "n_isects": n_isects},
        "pairs": pc,
        "stages_ms": {k: round(v[0], 4) for k, v in stages.items()},
        "eager_path": {"value": round(1e3 / eager_ms, 3), "unit": "iters/s", "ms_per_step": round(eager_ms, 3),
                       "description": "the same fused step issued eagerly through the autograd wrapper (per "
                                      "step: the input and adjoint value checks and the overflow check read "
                                      "back to the host)"},
        "plain_path": None if plain is None else {
            "value": round(1e3 / plain[0], 3), "unit": "iters/s", "ms_per_step": round(plain[0], 3),
            "speedup_of_value": round(plain[0] / ms, 3),
            "description": "same step through one HIP operator per stage (staged reference op sequence, "
                           "host-side intersection sizing)",
            "stages_ms": {k: round(v[0], 4) for k, v in plain[1].items()}},
        "voxelize_128": voxel,
        "fit_latency": latency,
        "parity_fp64": parity,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


def batched_main(args):
    """32 uneven reconstruction jobs (N log-uniform in [10k, 400k], C in
    {16, 32, 64, 128}, side in {128, 256, 512}), fused fwd+MSE+bwd+Adam
    (lr 1e-2) of every job per batched step from one hipGraph, kernel
    configuration per job from the autotuner; compared with one step of every
    job through the plain operator path with torch.optim.Adam."""
    import torch

    from r2splat_amd import autograd as ag
    from r2splat_amd import explicit, workloads
    from r2splat_amd.batched import Autotuner, BatchedFit
    from r2splat_amd.train import AdamConfig

    device = torch.device("cuda:0")
    specs = workloads.config3_specs(args.jobs, seed=0)
    t0 = time.perf_counter()
    jobs = [workloads.fit_job(n, c, side, seed, device=device) for n, c, side, seed in specs]
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    tuner = Autotuner(repeats=2)
    t0 = time.perf_counter()
    batch = BatchedFit(jobs, adam=AdamConfig(lr=1e-2, min_scale=1e-6), n_streams=args.streams, tuner=tuner,
                       packed=not args.no_packed)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    batch.step(1)
    batch.sync()
    losses0 = batch.losses()
    t0 = time.perf_counter()
    batch.step(args.steps)
    applied = batch.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    losses = batch.losses()
    n_isects = [int(f.counters[0].item()) for f in batch.fits]

    plain_ms = None
    if args.plain_steps > 0:
        explicit.BACKEND = "ops"
        total = 0.0
        for job in jobs:
            params = [job[k].detach().clone().requires_grad_() for k in ("means", "quats", "scales", "densities")]
            opt = torch.optim.Adam(params, lr=1e-2)

            def pstep():
                img = ag.rasterize(*params, job["viewmats"], job["Ks"], tuple(job["target"].shape[-2:]),
                                   recompute=False)
                loss = ((img - job["target"]) ** 2).mean()
                opt.zero_grad(set_to_none=True)
                loss.backward()
                opt.step()
                with torch.no_grad():  # the same domain clamp as the batched step
                    params[2].clamp_(min=1e-6)
                    params[3].clamp_(min=0.0)

            # every timed step starts from the job's initial parameters (restored outside the timed
            # region), so each job's plain step sees the same inputs as its first batched step
            initial = [p.detach().clone() for p in params]
            for k in range(args.plain_steps + 1):
                with torch.no_grad():
                    for p, p0 in zip(params, initial):
                        p.copy_(p0)
                torch.cuda.synchronize()
                s0 = time.perf_counter()
                pstep()
                torch.cuda.synchronize()
                if k > 0:
                    total += (time.perf_counter() - s0) / args.plain_steps
        plain_ms = total * 1e3
        explicit.BACKEND = "fused"

    configs = {}
    for f in batch.fits:
        key = f"tile{f.tile_size}/fwd{f.fwd_waves or 8}/bwd{f.bwd_waves or 8}"
        configs[key] = configs.get(key, 0) + 1
    # prepare + render + raster VJP of every job at its tuned configuration
    # against the default one (tile 16, 8 waves each), from the tuner's timings
    tuned_ms = sum(r.ms for r in batch.tune_records)
    default_ms = sum(r.timings[(16, "prepare", 0)] + r.timings[(16, "render", 8)] + r.timings[(16, "raster_vjp", 8)]
                     for r in batch.tune_records)
    result = {
        "metric": "batched fwd+MSE+bwd+Adam steps/s (32 uneven jobs, one hipGraph per step)",
        "value": round(1e3 / ms, 3),
        "unit": "batched steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (per-job seeded Gaussian sets and orbit cameras; targets from second seeded sets)",
        "config": {"workload": f"batched: {len(jobs)} jobs, N log-uniform [10k, 400k], C in {{16,32,64,128}}, "
                               f"side in {{128,256,512}}, Adam lr 1e-2", "streams": args.streams,
                   "packed_raster_launches": batch.packed,
                   "jobs": [list(sp[:3]) for sp in specs]},
        "job_steps_per_s": round(len(jobs) * 1e3 / ms, 2),
        "applied_steps": applied,
        "intersections_per_step": int(sum(n_isects)),
        "resident_gb": round(batch.resident_bytes() / 1e9, 2),
        "tuned_configs": configs,
        "job_configs": [[int(f.N), int(f.C), int(f.H), f.tile_size, f.fwd_waves or 8, f.bwd_waves or 8]
                        for f in batch.fits],
        "tuning": {"tuned_ms": round(tuned_ms, 2), "default_ms": round(default_ms, 2),
                   "description": "prepare + render + raster VJP of every job once, at its tuned configuration and "
                                  "at tile 16 / 8 waves (the tuner's timings)"},
        "setup_s": {"generate": round(t_gen, 2), "tune_and_build": round(t_tune, 2)},
        "mse_first_last": [round(sum(losses0) / len(losses0), 6), round(sum(losses) / len(losses), 6)],
        "plain_path": None if plain_ms is None else {
            "ms_per_batched_step": round(plain_ms, 2), "speedup_of_value": round(plain_ms / ms, 3),
            "description": "each job stepped alone through the autograd wrapper on the plain operator path "
                           "with torch.optim.Adam and the same clamp"},
    }
    print(json.dumps(result), flush=True)


def config2_main(args):
    """N = 500,000 Beta primitives with uncertainty, 256 views of 512^2:
    rasterize_fwd returning (mean, variance) + rasterize_bwd with adjoints
    N(0, 1e-3) / N(0, 1e-4), for detach_variance False and True; and the same
    primitives voxelized to 256^3 with the isotropic TV loss (weight 1e-3),
    forward and backward. Steps are timed eagerly (host-side sizing included)."""
    import torch

    import r2splat_amd
    from r2splat_amd import profiling, workloads
    from r2splat_amd import hip as ops

    device = torch.device("cuda:0")
    sc = workloads.config2(device=device)
    g = torch.Generator(device=device).manual_seed(5)
    v = torch.randn((256, 512, 512), generator=g, device=device) * 1e-3
    vv = torch.randn((256, 512, 512), generator=g, device=device) * 1e-4
    one = torch.ones((1,), device=device)

    def raster_step(detach):
        out, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size, beta=sc.beta,
                                             uncertainty=sc.uncertainty, detach_variance=detach)
        return r2splat_amd.rasterize_bwd(ctx, v, vv)

    def voxel_step():
        vol, ctx = r2splat_amd.voxelize_fwd(*sc.params(), (256, 256, 256), volume_extent=(1.0, 1.0, 1.0),
                                            beta=sc.beta)
        volume = vol[None]
        ops.volume_loss_iso_fwd(volume, 1e-3)
        gvol = ops.volume_loss_iso_bwd(volume, one, 1e-3)
        return r2splat_amd.voxelize_bwd(ctx, gvol[0].contiguous())

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        profiling.reset()
        profiling.enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        stages = {k: round(v_[0], 4) for k, v_ in profiling.collect().items()}
        profiling.enable(False)
        return ms, stages

    steps = max(1, min(args.steps, 10))
    r_ms, r_stages = timed(lambda: raster_step(False), steps, 1)
    d_ms, _ = timed(lambda: raster_step(True), steps, 1)
    v_ms, v_stages = timed(voxel_step, steps, 1)
    result = {
        "metric": "Beta+uncertainty rasterize fwd+bwd iters/s (500k Beta primitives, 256x512^2 views)",
        "value": round(1e3 / r_ms, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": round(r_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Beta primitives with uncertainty, orbit cameras, Gaussian adjoints)",
        "config": {"workload": "config2: N=500000 Beta primitives (beta log-uniform [0.5,4], uncertainty U[0,0.1], "
                               "scales [0.003,0.03]), C=256 views of 512x512, (mean, variance) fwd+bwd; voxels 256^3 "
                               "with isotropic TV 1e-3"},
        "stages_ms": r_stages,
        "detach_variance_ms_per_step": round(d_ms, 3),
        "voxel_tv_256": {"ms_per_step": round(v_ms, 3), "stages_ms": v_stages},
    }
    print(json.dumps(result), flush=True)


def serving_main(args):
    """Fit requests arriving as a Poisson stream (rate x duration; N log-uniform
    in [10k, 400k], C in {16, 32, 64, 128}, side in {128, 256, 512}, Gaussian
    or Beta with p = 0.5, `serve-steps` Adam iterations each) served by the
    continuous-batching server against the device's HBM; time-to-finished-fit
    p50/p99 from arrival and served iterations/s. A few requests are also
    fitted on the plain per-job path (explicit API over the staged operators,
    torch.optim.Adam) and compared, parameter for parameter, with the server's
    captured step fitted alone over their first iterations."""
    import torch

    from r2splat_amd import explicit
    from r2splat_amd.serving import ExplicitFit, FitServer, build_job, job_config, poisson_requests
    from r2splat_amd.train import AdamConfig, GaussianFit

    device = torch.device("cuda:0")
    adam = AdamConfig(lr=1e-2, min_scale=1e-6)
    reqs = poisson_requests(args.rate, args.duration, seed=0, steps=args.serve_steps)
    server = FitServer(device, n_streams=args.streams, adam=adam, max_active=args.max_active or None,
                       active_ms=args.active_ms or None)

    def progress(stats, live, queued):
        print(f"[serving] {time.perf_counter() - t0:.0f}s finished {len(stats.finished)}/{len(reqs)} live {live} "
              f"queued {queued} job-steps {stats.job_steps}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    stats = server.run(reqs, progress=progress, progress_every=20.0)
    lat = stats.latencies()
    final = {f[0]: f[4] for f in stats.finished}

    # the plain per-job path (explicit API over the staged operators, torch Adam) against the same
    # request fitted alone by the server's captured step, over the first few iterations: 200 steps of
    # Adam at lr 1e-2 amplify rounding into different (equally valid) fits, the first steps do not
    plain = []
    explicit.BACKEND = "ops"
    for req in reqs[:max(0, args.plain_jobs)]:
        job = build_job(req, device)
        fit = ExplicitFit(job, adam)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        for _ in range(PLAIN_CHECK_STEPS):
            fit.step()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - s0) * 1e3 / PLAIN_CHECK_STEPS
        explicit.BACKEND = "fused"
        cfg = job_config(req)
        alone = GaussianFit(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"],
                            job["target"], adam=adam, beta=job["beta"], tile_size=cfg.tile_size,
                            fwd_waves=cfg.fwd_waves, bwd_waves=cfg.bwd_waves)
        alone.capture()
        alone.step(PLAIN_CHECK_STEPS)
        alone.sync()
        explicit.BACKEND = "ops"
        ref = torch.cat([q.detach().reshape(-1) for q in fit.params])
        err = float((alone.params.double() - ref.double()).abs().max() / ref.double().abs().max())
        plain.append({"rid": req.rid, "n": req.n, "views": req.views, "side": req.side, "beta": req.beta,
                      "steps_compared": PLAIN_CHECK_STEPS, "param_rel_err": float(f"{err:.3e}"),
                      "final_mse_served": round(final[req.rid], 8),
                      "plain_ms_per_iter": round(plain_ms, 3)})
        del alone, fit, job
        print(f"[serving] plain path request {req.rid} done", file=sys.stderr, flush=True)
    explicit.BACKEND = "fused"

    result = {
        "metric": "time-to-finished-fit p50/p99 and served iters/s (Poisson fit requests, continuous batching)",
        "value": round(stats.job_steps / stats.wall, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": stats.job_steps,
        "warmup": 0,
        "ms_per_step": round(stats.wall * 1e3 / max(stats.job_steps, 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (per-request seeded Gaussian/Beta sets and orbit cameras; targets from second seeded sets)",
        "config": {"workload": f"serving: Poisson {args.rate} req/s for {args.duration} s, N log-uniform [10k, 400k], "
                               f"C in {{16,32,64,128}}, side in {{128,256,512}}, Gaussian or Beta (p=0.5), "
                               f"{args.serve_steps} Adam iterations (lr 1e-2) each", "streams": args.streams,
                   "max_active": args.max_active or None, "active_ms": args.active_ms or None,
                   "hbm_budget_gb": round(server.budget / 1e9, 1)},
        "requests": len(reqs),
        "finished": len(stats.finished),
        "beta_requests": sum(r.beta for r in reqs),
        "latency_s": {"p50": round(stats.percentile(50), 2), "p99": round(stats.percentile(99), 2),
                      "min": round(lat[0], 2) if lat else None, "max": round(lat[-1], 2) if lat else None},
        "wall_s": round(stats.wall, 1),
        "ticks": stats.ticks,
        "peak_live_jobs": stats.peak_live,
        "peak_resident_gb": round(stats.peak_bytes / 1e9, 2),
        "plain_path": plain,
    }
    print(json.dumps(result), flush=True)


def parity_vs_fp64(sc, target):
    """Max relative error (max |a - b| / max |b|) of the benchmarked float32 step
    (image and every gradient, MSE adjoint against the target) against float64
    on the same inputs: over the pairs the float32 kernels kept (the float32
    arithmetic) and against the float64 operator path itself (which also
    counts pairs at the 1e-5 cut-off that the two precisions decide
    differently, and footprints whose radius rounds across an integer), with
    the number of each. Evaluated once, outside the timed region."""
    import torch

    import r2splat_amd
    from r2splat_amd import parity

    C, N = sc.viewmats.shape[0], sc.means.shape[0]
    with torch.no_grad():
        img, ctx = r2splat_amd.rasterize_fwd(*sc.params(), sc.viewmats, sc.Ks, sc.image_size)
        table = parity.pair_table(ctx.buffers["fused"], C, N)
        v = (2.0 * (img - target) / img.numel()).contiguous()
        grads = r2splat_amd.rasterize_bwd(ctx, v)
        sc64 = parity.fp64_scene(*sc.params(), sc.viewmats, sc.Ks, sc.image_size)
        kept = parity.evaluate(table, sc64, "kept", v_image=v)
        coverage = parity.coverage_diff(table, parity.fp64_table(sc64), C, N)
        img64 = sc64.image
        g64 = r2splat_amd.rasterize_bwd(sc64.ctx, v.double())

    def rel(a, b):
        a, b = a.double(), b.double()
        return float(f"{float((a - b).abs().max() / b.abs().max()):.3e}")

    out = {"image": {"kept_pairs": rel(img, kept.image), "float64_reference": rel(img, img64)}}
    for name in ("means", "quats", "scales", "densities"):
        out[name] = {"kept_pairs": rel(getattr(grads, name), kept.grads[name]),
                     "float64_reference": rel(getattr(grads, name), getattr(g64, name))}
    out["pairs"] = kept.pairs
    out["cut_off_flips"] = {"forward": kept.flips_fwd, "backward": kept.flips_bwd}
    out["coverage_differences"] = coverage
    out["bounds"] = "images 1e-5, gradients 1e-4 (max |a - b| / max |b|)"
    del table, sc64, kept, g64
    torch.cuda.empty_cache()
    return out


def fit_latency(rate, duration, steps, streams, max_active=None, active_ms=None):
    """p50 / p99 time-to-finished-fit of Poisson fit requests (the serving
    workload's request mix) over a short arrival window, served by the
    continuous-batching server on this GPU."""
    import torch

    from r2splat_amd.serving import FitServer, poisson_requests
    from r2splat_amd.train import AdamConfig

    reqs = poisson_requests(rate, duration, seed=0, steps=steps)
    server = FitServer(torch.device("cuda:0"), n_streams=streams, adam=AdamConfig(lr=1e-2, min_scale=1e-6),
                       max_active=max_active, active_ms=active_ms)
    stats = server.run(reqs)
    del server
    torch.cuda.empty_cache()
    return {"p50_s": round(stats.percentile(50), 3), "p99_s": round(stats.percentile(99), 3),
            "served_iters_per_s": round(stats.job_steps / stats.wall, 2), "requests": len(reqs),
            "finished": len(stats.finished), "wall_s": round(stats.wall, 2),
            "window": f"Poisson {rate:g} req/s over {duration:g} s (seed 0): N log-uniform [10k, 400k], C in "
                      f"{{16,32,64,128}}, side in {{128,256,512}}, Gaussian or Beta (p=0.5), {steps} Adam "
                      f"iterations (lr 1e-2) each; latency from arrival to finished fit"}


def graph_timing(step, steps, warmup):
    """Captures one step (autograd wrapper, fused path) into a hipGraph and
    times `steps` replays; the intersection capacity is the one sized by the
    eager warm-up, and the overflow flag is checked after the timed region."""
    import torch

    from r2splat_amd import explicit, fused

    explicit.BACKEND = "fused"
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(max(1, warmup)):
            step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for _ in range(max(1, warmup)):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    if fused.LAST_COUNTERS is not None and int(fused.LAST_COUNTERS[1].item()) != 0:
        raise RuntimeError("intersection capacity overflowed inside the captured step")
    return ms


def voxel_timing(sc, steps):
    """voxelize fwd+bwd of the same Gaussians to 128^3, extent (1, 1, 1), centre 0."""
    import torch

    from r2splat_amd import autograd as ag
    from r2splat_amd import profiling

    params = [t.detach().requires_grad_() for t in sc.params()]
    shape = (128, 128, 128)
    w = torch.randn(shape, generator=torch.Generator(device=sc.means.device).manual_seed(3),
                    device=sc.means.device)

    def vstep():
        vol = ag.voxelize(*params, shape, volume_extent=(1.0, 1.0, 1.0), recompute=False)
        return torch.autograd.grad(vol, params, w)

    vstep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        vstep()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    # stage times from a second pass with stage events (which add host work)
    profiling.reset()
    profiling.enable(True)
    for _ in range(steps):
        vstep()
    torch.cuda.synchronize()
    profiling.enable(False)
    stages = {k: round(v[0], 4) for k, v in profiling.collect().items()}
    # intersections (primitive, 8^3 tile) of this volume: each kernel evaluates
    # every voxel of each intersected tile (512 voxel-primitive pairs)
    from r2splat_amd import explicit

    with torch.no_grad():
        _, ctx = explicit.voxelize_fwd(*sc.params(), shape, volume_extent=(1.0, 1.0, 1.0))
    n = int(ctx.buffers["isect_ids"].numel())
    del ctx
    pairs = n * 512
    rates = {name: round(pairs / (stages[name] * 1e-3) / 1e9, 1) for name in ("voxelize_fwd", "voxelize_bwd")
             if stages.get(name)}
    return {"ms_per_step": round(ms, 3), "iters_per_s": round(1e3 / ms, 3), "stages_ms": stages,
            "intersections": n, "tile_pairs": pairs, "tile_pairs_per_s_e9": rates}


def _popcount(words):
    """Set bits of an int32 tensor (SWAR in int64)."""
    x = words.long() & 0xFFFFFFFF
    x = x - ((x >> 1) & 0x55555555)
    x = (x & 0x33333333) + ((x >> 2) & 0x33333333)
    x = (x + (x >> 4)) & 0x0F0F0F0F
    return int((((x * 0x01010101) & 0xFFFFFFFF) >> 24).sum().item())


def pair_counts(sc):
    """Pixel-primitive pairs of the workload: in the intersected tiles
    (intersections x 256) and those the fused forward and backward keep at the
    cut-off (their keep masks, r2s_fused_keep_masks; 256^2 images have no
    pixels past the edge of a 16-pixel tile)."""
    import torch

    from r2splat_amd import fused
    from r2splat_amd.hip import _backend as be

    with torch.no_grad():
        fwd = fused.prepare(*sc.params(), sc.viewmats, sc.Ks, sc.width, sc.height, 1e-8, 3.0)
        n, T = fwd.n_isects, fwd.tile_size
        words = T * T // 32
        mf = torch.empty((max(n, 1), words), dtype=torch.int32, device=sc.means.device)
        mb = torch.empty_like(mf)
        be.call("r2s_fused_keep_masks", be.ptr(fwd.packed), be.ptr(fwd.offsets), be.ptr(fwd.counters), fwd.capacity,
                sc.viewmats.shape[0], sc.width, sc.height, T, fwd.kind, be.ptr(mf), be.ptr(mb),
                be.stream_of(sc.means.device))
        out = {"intersections": n, "list_pairs": n * T * T, "kept_fwd": _popcount(mf[:n]),
               "kept_bwd": _popcount(mb[:n])}
        del fwd, mf, mb
    return out


def cpu_baseline(sc, target, views, min_seconds):
    import numpy as np

    from oracle import cpu_baseline as cb

    C = sc.viewmats.shape[0]
    views = max(1, min(views, C))
    cores = cb.host_cores()
    # the box allots a share of its cores to each GPU and says so in OMP_NUM_THREADS
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(cores, share) if share > 0 else cores
    params = [t.detach().cpu().numpy() for t in sc.params()]
    vm = sc.viewmats[:views].cpu().numpy()
    K = sc.Ks[:views].cpu().numpy()
    tg = target[:views].cpu().numpy().astype(np.float32)
    reps, t0 = 0, time.perf_counter()
    while True:
        cb.iteration(*params, vm, K, sc.height, sc.width, tg, threads=threads)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    per_iter = dt / reps * C / views
    return {"value": round(1.0 / per_iter, 5), "unit": "iters/s", "cores": threads, "kind": "port",
            "host": {"cpu_model": _cpu_model(), "cores_visible": cores,
                     "omp_num_threads": os.environ.get("OMP_NUM_THREADS")},
            "sample": f"{reps} fwd+MSE+bwd iterations over {views} of {C} views x all {sc.means.shape[0]} "
                      f"Gaussians ({dt:.1f} s), scaled by {C}/{views}; {threads} OpenMP threads (all the "
                      f"cores this process may use: the box's per-GPU CPU share)"}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()

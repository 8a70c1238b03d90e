"""Batched fits: many independent reconstructions of uneven size stepped
together from one hipGraph, with per-job kernel configurations chosen by an
autotuner.

``BatchedFit`` places each job (a ``GaussianFit``: its own Gaussians, views,
image size and Adam state) on one of a few streams, longest jobs first onto
the least-loaded stream, and captures one step of every job (fused forward,
MSE, backward, Adam) into a single graph whose streams fork from and join
one capture stream, so a whole batched step costs one graph launch.

With ``packed`` (the default) the two raster stages of all jobs run as packed
launches: one render and one raster-VJP launch per group of jobs that share a
tile edge, wave count and record kind, over the concatenation of those jobs'
tiles (``r2s_fused_render_packed`` / ``r2s_fused_raster_vjp_packed``), so
small jobs fill the GPU together instead of one after another. The graph then
runs every job's prepare on its stream, the packed renders, every job's loss,
the packed raster VJPs, and every job's parameter VJP and Adam update. Each job
keeps its own ``Workspace`` (the packed launches read all jobs' intersection
lists at once). Without ``packed`` the jobs run their whole steps on their
streams and jobs on the same stream share one workspace, sized for the
largest of them.

``Autotuner`` picks the tile edge (8, 16, 32) and the waves per tile of the
forward and backward raster kernels (4, 8, 16) for a job by timing the
candidates on that job's own data; results are cached by (views, image size,
log2 N). Every configuration computes the same function (the tile edge only
changes which pixels share a list, the waves only the forward's summation
order).

A job whose intersection count outgrows its capacity skips its own update on
the device (see ``GaussianFit``); ``sync()`` grows that job's capacity,
replays its skipped steps on its own and re-captures the batch, so every
job's sequence of applied updates equals that of the job fitted alone.

The reference fits one scene per call (r2splat/explicit.py); this layer is
the batched-fit capability built on the same operators.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from . import fused
from .hip import _backend as be
from .hip._backend import call, ptr
from .train import AdamConfig, GaussianFit, Workspace


@dataclass(frozen=True)
class KernelConfig:
    tile_size: int = 16
    fwd_waves: int = 8
    bwd_waves: int = 8


def heuristic_config(n: int, views: int, height: int, width: int, beta: bool = False) -> KernelConfig:
    """A kernel configuration from a job's shape alone (``fused.config_for_shape``)."""
    return KernelConfig(*fused.config_for_shape(n, views, height, width, beta))


@dataclass
class TuneRecord:
    config: KernelConfig
    ms: float          # prepare + render + raster VJP of the chosen configuration
    timings: dict      # (tile, kind, waves) -> ms


class Autotuner:
    """Times tile edges and waves per tile on a job's own data and keeps the fastest."""

    TILES = (8, 16, 32)
    WAVES = (4, 8, 16)

    def __init__(self, repeats: int = 2, tiles: Sequence[int] = TILES, waves: Sequence[int] = WAVES):
        self.repeats = max(1, int(repeats))
        self.tiles = tuple(tiles)
        self.waves = tuple(waves)
        self.cache: dict = {}

    @staticmethod
    def key(N: int, C: int, H: int, W: int) -> tuple:
        return (C, H, W, int(round(math.log2(max(N, 1)))))

    def _time(self, fn) -> float:
        fn()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(self.repeats):
            fn()
        end.record()
        end.synchronize()
        return start.elapsed_time(end) / self.repeats

    def tune(self, means, quats, scales, densities, viewmats, Ks, height: int, width: int, *, eps: float = 1e-8,
             radius_clip: float = 3.0, beta=None) -> TuneRecord:
        C, N = int(viewmats.shape[0]), int(means.shape[0])
        kind = int(beta is not None)
        key = self.key(N, C, height, width) + (kind,)
        if key in self.cache:
            return self.cache[key]
        device = means.device
        v_image = torch.randn((C, height, width), device=device, dtype=torch.float32,
                              generator=torch.Generator(device=device).manual_seed(0)) * 1e-6
        timings, best = {}, None
        factors = fused.beta_factors(beta) if kind else None
        for tile in self.tiles:
            prep = lambda: fused.prepare(means, quats, scales, densities, viewmats, Ks, width, height, eps,  # noqa
                                         radius_clip, tile_size=tile, beta=beta, factors=factors)
            f = prep()
            t_prep = self._time(prep)
            grads = torch.empty((f.capacity * 8,), dtype=torch.float32, device=device)
            stream = be.stream_of(device)
            fwd, bwd = {}, {}
            for w in self.waves:
                fwd[w] = self._time(lambda: fused.render(f, C, device, w))
                bwd[w] = self._time(lambda: call(
                    "r2s_fused_raster_vjp", ptr(f.packed), ptr(f.offsets), ptr(f.tile_order), ptr(f.counters),
                    f.capacity, C, width, height, tile, w, kind, ptr(v_image), None, 0, ptr(grads), stream))
                timings[(tile, "render", w)] = fwd[w]
                timings[(tile, "raster_vjp", w)] = bwd[w]
            timings[(tile, "prepare", 0)] = t_prep
            wf, wb = min(fwd, key=fwd.get), min(bwd, key=bwd.get)
            total = t_prep + fwd[wf] + bwd[wb]
            if best is None or total < best.ms:
                best = TuneRecord(KernelConfig(tile, wf, wb), total, timings)
            del f, grads
        best.timings = timings
        self.cache[key] = best
        return best


class _PackedGroup:
    """Jobs that share a tile edge, wave count and record kind: the job table
    of their packed raster launch (host copy and device copy)."""

    def __init__(self, fits, tile_size: int, waves: int, kind: int, device):
        self.fits = list(fits)
        self.tile_size, self.waves, self.kind = tile_size, waves, kind
        table = (be.RasterJob * len(self.fits))()
        first = 0
        for entry, fit in zip(table, self.fits):
            for name, value in fit.raster_job().items():
                setattr(entry, name, value)
            entry.first_tile = first
            first += fit.C * (-(-fit.H // tile_size)) * (-(-fit.W // tile_size))
        self.tiles = first
        self.host = table
        raw = torch.frombuffer(bytearray(ctypes.string_at(ctypes.addressof(table), ctypes.sizeof(table))),
                               dtype=torch.uint8)
        self.table = raw.to(device)  # the copy the kernels read (kept alive with the graph)

    def issue(self, name: str, stream):
        call(name, ctypes.addressof(self.host), ptr(self.table), len(self.fits), self.tile_size, self.waves,
             self.kind, stream)


def _groups(fits, waves_of, device):
    """Packed groups by (tile edge, waves, kind); within a group the jobs with
    the most primitives per tile of a view come first, so each XCD (which takes
    its band of every job in table order) ends on the cheapest tiles."""
    groups = {}
    for fit in fits:
        key = (fit.tile_size, waves_of(fit) or 8, fit.kind)
        groups.setdefault(key, []).append(fit)
    per_tile = lambda f: f.N / ((-(-f.H // f.tile_size)) * (-(-f.W // f.tile_size)))  # noqa: E731
    return [_PackedGroup(sorted(members, key=per_tile, reverse=True), *key, device)
            for key, members in sorted(groups.items())]


def _clip(job, radius_clip):
    """The explicit API's default footprint clip: 3 sigma (Gaussian), 1 (Beta support)."""
    if radius_clip is not None:
        return radius_clip
    return 3.0 if job.get("beta") is None else 1.0


class BatchedFit:
    """Steps many independent ``GaussianFit`` jobs together from one hipGraph.

    ``jobs``: sequence of dicts with means, quats, scales, densities, viewmats,
    Ks, target (float32, on one device) and optionally ``beta`` (Beta
    primitives), a ``config`` (KernelConfig) and ``capacity``. Gaussian and
    Beta jobs mix freely in one batch. A job without a config takes the
    tuner's (when a tuner is given), else the configuration of its shape
    (``heuristic_config``).
    """

    def __init__(self, jobs, *, adam: Optional[AdamConfig] = AdamConfig(), n_streams: int = 4,
                 tuner: Optional[Autotuner] = None, eps: float = 1e-8, radius_clip: Optional[float] = None,
                 packed: bool = True):
        if not jobs:
            raise ValueError("no jobs")
        self.device = jobs[0]["means"].device
        self.packed = bool(packed)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(max(1, n_streams))]
        stream_ws = [Workspace(self.device) for _ in self.streams]
        self.workspaces = [] if self.packed else stream_ws
        self.tune_records = []
        load = [0.0] * len(self.streams)
        order = sorted(range(len(jobs)), key=lambda i: -self._cost(jobs[i]))
        self.fits: list = [None] * len(jobs)
        for i in order:
            job = jobs[i]
            cfg = job.get("config")
            if cfg is None and tuner is not None:
                rec = tuner.tune(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"],
                                 job["Ks"], int(job["target"].shape[-2]), int(job["target"].shape[-1]), eps=eps,
                                 radius_clip=_clip(job, radius_clip), beta=job.get("beta"))
                cfg = rec.config
                self.tune_records.append(rec)
            if cfg is None:  # the configuration of the job's shape
                C, H, W = job["target"].shape
                cfg = heuristic_config(int(job["means"].shape[0]), int(C), int(H), int(W), job.get("beta") is not None)
            s = min(range(len(self.streams)), key=lambda k: load[k])
            load[s] += self._cost(job)
            ws = Workspace(self.device) if self.packed else stream_ws[s]
            if self.packed:
                self.workspaces.append(ws)
            self.fits[i] = GaussianFit(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"],
                                       job["Ks"], job["target"], eps=eps, radius_clip=_clip(job, radius_clip),
                                       adam=adam, capacity=job.get("capacity"), stream=self.streams[s],
                                       tile_size=cfg.tile_size, fwd_waves=cfg.fwd_waves, bwd_waves=cfg.bwd_waves,
                                       workspace=ws, beta=job.get("beta"))
        self._render_groups = self._vjp_groups = None
        self.graph = None
        self._versions = None
        self.steps_issued = 0

    @staticmethod
    def _cost(job) -> float:
        C, H, W = job["target"].shape
        return float(job["means"].shape[0]) * C * math.sqrt(H * W)

    def _state(self) -> list:
        """What a captured graph bakes in: the workspace buffers (their versions)
        and each job's capacity (a job can grow its capacity inside a shared
        workspace that is already large enough)."""
        return [ws.version for ws in self.workspaces] + [f.capacity for f in self.fits]

    def _stale(self) -> bool:
        return self.graph is None or self._versions != self._state()

    def _per_job(self, main, stage):
        """Fork to the job streams, run `stage(fit)` for each job on its stream, join."""
        for s in self.streams:
            s.wait_stream(main)
        for fit in self.fits:
            with torch.cuda.stream(fit.stream):
                stage(fit)
        for s in self.streams:
            main.wait_stream(s)

    def _per_group(self, main, groups, name):
        """Fork, one packed launch per group (groups dealt over the streams, so
        one group's tail overlaps the next group), join."""
        for s in self.streams:
            s.wait_stream(main)
        for i, group in enumerate(groups):
            s = self.streams[i % len(self.streams)]
            group.issue(name, s.cuda_stream)
        for s in self.streams:
            main.wait_stream(s)

    def capture(self):
        """One graph: fork to the job streams, each job's step in order, join
        (packed: each job's prepare, the packed renders, each job's loss, the
        packed raster VJPs, each job's update)."""
        main = torch.cuda.Stream(self.device)
        main.wait_stream(torch.cuda.current_stream(self.device))
        for s in self.streams:  # buffers written before capture are visible to every job stream
            s.wait_stream(torch.cuda.current_stream(self.device))
        if self.packed:  # the job tables hold the current buffers
            self._render_groups = _groups(self.fits, lambda f: f.fwd_waves, self.device)
            self._vjp_groups = _groups(self.fits, lambda f: f.bwd_waves, self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(main):
            with torch.cuda.graph(g, stream=main):
                if not self.packed:
                    self._per_job(main, lambda fit: fit._issue())
                else:
                    self._per_job(main, lambda fit: fit._issue_prepare())
                    self._per_group(main, self._render_groups, "r2s_fused_render_packed")
                    self._per_job(main, lambda fit: fit._issue_loss())
                    self._per_group(main, self._vjp_groups, "r2s_fused_raster_vjp_packed")
                    self._per_job(main, lambda fit: fit._issue_update())
        torch.cuda.current_stream(self.device).wait_stream(main)
        self.graph = g
        self._versions = self._state()
        return g

    def step(self, n: int = 1):
        """Issues n batched steps (graph replays) without synchronising."""
        if self._stale():
            self.capture()
        for _ in range(n):
            self.graph.replay()
        self.steps_issued += n

    def sync(self) -> int:
        """Waits, then repairs jobs that overflowed their capacity (grow,
        replay their skipped steps alone) until every job has applied the same
        number of updates. Returns that number."""
        torch.cuda.synchronize(self.device)
        while True:
            pending = [(fit, fit.pending_skips()) for fit in self.fits]
            pending = [(fit, k) for fit, k in pending if k > 0]
            if not pending:
                break
            for fit, k in pending:
                fit.regrow()
                with torch.cuda.stream(fit.stream):
                    fit.replay_skipped(k, graph=False)
            torch.cuda.synchronize(self.device)
        applied = {int(f.adam_state[0].item()) for f in self.fits} if self.fits[0].adam is not None else {
            self.steps_issued}
        if len(applied) != 1:
            raise RuntimeError(f"jobs applied different step counts: {sorted(applied)}")
        return applied.pop()

    def resident_bytes(self) -> int:
        return sum(f.resident_bytes() for f in self.fits) + sum(ws.nbytes() for ws in self.workspaces)

    def losses(self) -> list:
        return [float(v) for v in torch.cat([f.loss for f in self.fits]).tolist()]


__all__ = ["Autotuner", "BatchedFit", "KernelConfig", "TuneRecord", "heuristic_config"]

"""Continuous-batching reconstruction server.

Requests (one reconstruction job each: N Gaussians or Beta primitives, C
views of side x side, a number of Adam iterations) arrive over time. The
server admits a queued request as soon as its HBM footprint fits the budget
(in arrival order, letting a later request that fits pass one that does
not), keeps every admitted job's parameters, optimizer state, target images
and intermediates resident on the device, and on every tick issues one
iteration of every live job: each job (Gaussian or Beta primitives)
replays its own captured hipGraph (fused fwd + MSE + bwd + Adam) on one of a
few streams, so jobs overlap on the device. A job leaves, and frees its
memory, when its last iteration has been applied; time-to-finished-fit is
measured from its arrival.

Updates are Adam followed by a clamp to the explicit API's domain (scales >=
``min_scale``, densities >= 0, beta >= ``min_beta``) so that the same job can
be stepped by the plain explicit path (``ExplicitFit``) for comparison. A
job's result does not depend on which other jobs share the device with it:
the fused path is bitwise deterministic for both primitive kinds.

The reference has no serving layer; this is the serving capability built on
the same operators.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import torch

from . import explicit, workloads
from .batched import KernelConfig, heuristic_config
from .train import AdamConfig, GaussianFit


@dataclass
class FitRequest:
    rid: int
    arrival: float          # seconds (or ticks, with a tick clock) from the start
    n: int
    views: int
    side: int
    seed: int
    beta: bool = False
    steps: int = 200


def poisson_requests(rate: float = 4.0, duration: float = 120.0, *, seed: int = 0, steps: int = 200,
                     p_beta: float = 0.5, n_range=(10_000, 400_000)) -> List[FitRequest]:
    """Poisson arrivals at `rate` per second for `duration` seconds; job sizes as
    in the batched workload (N log-uniform, C in {16, 32, 64, 128}, side in
    {128, 256, 512}); Beta with probability `p_beta`."""
    g = torch.Generator().manual_seed(seed)
    out, t, rid = [], 0.0, 0
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    while True:
        t += float(torch.empty(()).exponential_(rate, generator=g))
        if t >= duration:
            return out
        n = int(round(math.exp(lo + float(torch.rand((), generator=g, dtype=torch.float64)) * (hi - lo))))
        c = (16, 32, 64, 128)[int(torch.randint(0, 4, (), generator=g))]
        side = (128, 256, 512)[int(torch.randint(0, 3, (), generator=g))]
        beta = bool(torch.rand((), generator=g) < p_beta)
        out.append(FitRequest(rid, t, n, c, side, 7000 + 2 * rid, beta, steps))
        rid += 1


def job_config(req: FitRequest) -> KernelConfig:
    """The kernel configuration the server gives a request's job (tile edge,
    waves per tile of the raster forward and backward), from its shape."""
    return heuristic_config(req.n, req.views, req.side, req.side, req.beta)


def build_job(req: FitRequest, device) -> dict:
    """Parameters (seed) and target images (second set, seed + 1) of a request."""
    beta_range = (0.5, 4.0) if req.beta else None
    sc = workloads.scene(req.n, req.views, req.side, req.side, scale_range=(0.005, 0.05), seed=req.seed,
                         device=device, beta_range=beta_range)
    tg = workloads.scene(req.n, req.views, req.side, req.side, scale_range=(0.005, 0.05), seed=req.seed + 1,
                         device=device, beta_range=beta_range)
    with torch.no_grad():
        target, _ = explicit.rasterize_fwd(*tg.params(), tg.viewmats, tg.Ks, tg.image_size, beta=tg.beta)
    return dict(means=sc.means, quats=sc.quats, scales=sc.scales, densities=sc.densities, beta=sc.beta,
                viewmats=sc.viewmats, Ks=sc.Ks, target=target)


class ExplicitFit:
    """One job stepped through the explicit API eagerly (any primitive kind):
    rasterize_fwd -> MSE -> rasterize_bwd -> torch Adam -> clamp. The
    un-captured, one-call-per-stage baseline of the server's steps."""

    def __init__(self, job: dict, adam: AdamConfig, stream=None):
        self.job = job
        self.names = [k for k in ("means", "quats", "scales", "densities", "beta") if job.get(k) is not None]
        self.params = [job[k].detach().clone() for k in self.names]
        self.opt = torch.optim.Adam(self.params, lr=adam.lr, betas=(adam.beta1, adam.beta2), eps=adam.eps,
                                    foreach=False)
        self.adam = adam
        self.stream = stream
        self.applied = 0
        self.loss = None

    def step(self):
        with torch.cuda.stream(self.stream) if self.stream is not None else torch.cuda.device(self.params[0].device):
            kw = dict(zip(self.names, self.params))
            tgt = self.job["target"]
            img, ctx = explicit.rasterize_fwd(kw["means"], kw["quats"], kw["scales"], kw["densities"],
                                              self.job["viewmats"], self.job["Ks"], tuple(tgt.shape[-2:]),
                                              beta=kw.get("beta"))
            diff = img - tgt
            self.loss = (diff * diff).mean()
            grads = explicit.rasterize_bwd(ctx, (2.0 / img.numel()) * diff)
            for p, name in zip(self.params, self.names):
                p.grad = getattr(grads, name)
            self.opt.step()
            with torch.no_grad():
                k = self.names.index("scales")
                self.params[k].clamp_(min=self.adam.min_scale or 1e-6)
                self.params[self.names.index("densities")].clamp_(min=0.0)
                if "beta" in self.names:
                    self.params[self.names.index("beta")].clamp_(min=self.adam.min_beta)
        self.applied += 1

    def resident_bytes(self) -> int:
        t = self.job["target"]
        return 4 * sum(p.numel() for p in self.params) + t.numel() * t.element_size()


@dataclass
class _Live:
    req: FitRequest
    fit: GaussianFit
    admitted: float
    issued: int = 0
    nbytes: int = 0
    stalled: bool = False   # its capacity could not grow yet (device memory); retried at each sync


@dataclass
class ServeStats:
    finished: list = field(default_factory=list)   # (rid, arrival, admitted, finished, final_loss, beta)
    results: dict = field(default_factory=dict)    # rid -> final parameters (when kept)
    ticks: int = 0
    job_steps: int = 0
    peak_bytes: int = 0
    peak_live: int = 0
    wall: float = 0.0

    def latencies(self) -> list:
        return sorted(f[3] - f[1] for f in self.finished)

    def percentile(self, q: float) -> float:
        lat = self.latencies()
        if not lat:
            return float("nan")
        k = min(len(lat) - 1, max(0, int(math.ceil(q / 100.0 * len(lat))) - 1))
        return lat[k]


class FitServer:
    def __init__(self, device="cuda", *, hbm_budget: Optional[int] = None, n_streams: int = 4,
                 adam: AdamConfig = AdamConfig(lr=1e-2, min_scale=1e-6), headroom: float = 1.15,
                 sync_every: int = 8, config: Optional[Callable[[FitRequest], KernelConfig]] = None,
                 max_active: Optional[int] = None, active_ms: Optional[float] = None):
        """``config`` maps a request to its job's kernel configuration
        (default ``job_config``: from the job's shape, no tuning runs).
        ``max_active`` caps the jobs that step on a tick to the earliest
        arrived ones; ``active_ms`` steps the earliest arrived jobs until their
        estimated iteration times (``estimate_ms``) add up to that many
        milliseconds (enough work to keep the GPU busy); None for both: every
        live job steps on every tick."""
        self.device = torch.device(device)
        self.config = config or job_config
        self.max_active = None if max_active is None else max(1, int(max_active))
        self.active_ms = None if active_ms is None else float(active_ms)
        if hbm_budget is None:
            hbm_budget = int(torch.cuda.get_device_properties(self.device).total_memory * 0.85)
        self.budget = int(hbm_budget)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(max(1, n_streams))]
        self.adam = adam
        self.headroom = headroom
        self.sync_every = max(1, int(sync_every))
        self.live: List[_Live] = []
        self._next_stream = 0

    @staticmethod
    def estimate_bytes(req: FitRequest) -> int:
        """Footprint before building: parameters/state, images, and ~100 B per
        expected intersection (records, sort scratch, partials) with the
        expected intersections per (view, primitive) growing with the side."""
        per = {128: 2.0, 256: 5.0, 512: 15.0}.get(req.side, 5.0 * (req.side / 256.0) ** 2)
        isects = req.n * req.views * per
        return int(16 * 11 * req.n + 12 * req.views * req.side * req.side + 100 * isects + 48 * req.views * req.n)

    @staticmethod
    def estimate_ms(req: FitRequest) -> float:
        """Rough GPU time of one iteration of a request's job on MI355X: a
        fixed launch cost, then per expected intersection and per pixel (the
        config 1 shape, 27.5M intersections and 4.2M pixels, takes ~3.5 ms)."""
        per = {128: 2.0, 256: 5.0, 512: 15.0}.get(req.side, 5.0 * (req.side / 256.0) ** 2)
        return 0.15 + 1.0e-7 * req.n * req.views * per + 1.5e-7 * req.views * req.side * req.side

    def used_bytes(self) -> int:
        """Current footprint of the live jobs (their workspaces grow when their
        intersection counts do)."""
        return sum(l.fit.memory_bytes() for l in self.live)

    def _admit(self, req: FitRequest, now: float) -> bool:
        if any(l.stalled for l in self.live):  # a live job is waiting for memory
            return False
        if self.used_bytes() + self.estimate_bytes(req) > self.budget:
            return False
        job = build_job(req, self.device)
        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        cfg = self.config(req)
        fit = GaussianFit(job["means"], job["quats"], job["scales"], job["densities"], job["viewmats"], job["Ks"],
                          job["target"], adam=self.adam, stream=stream, headroom=self.headroom, beta=job["beta"],
                          tile_size=cfg.tile_size, fwd_waves=cfg.fwd_waves, bwd_waves=cfg.bwd_waves)
        fit.capture()
        self.live.append(_Live(req, fit, now, 0, fit.memory_bytes()))
        return True

    def _sync(self, l: _Live):
        """Synchronises one job; when growing its capacity runs out of device
        memory the job stalls (its overflowed steps stay skipped and are
        replayed by a later sync) and the allocator's cache is released, so
        that jobs finishing meanwhile make room."""
        try:
            l.fit.sync()
            l.stalled = False
        except torch.OutOfMemoryError:
            l.stalled = True
            torch.cuda.empty_cache()

    def run(self, requests: List[FitRequest], clock=None, max_wall: Optional[float] = None,
            keep_results: bool = False, progress: Optional[Callable[[ServeStats, int, int], None]] = None,
            progress_every: float = 30.0) -> ServeStats:
        """Serves `requests` (sorted by arrival) until all finished. `clock`
        returns the current time in the requests' unit (default: wall seconds
        since the call; "ticks": the number of ticks served so far);
        ``max_wall`` stops early (unfinished jobs dropped); ``keep_results``
        keeps each job's final parameters in ``stats.results``; ``progress``
        (stats, live jobs, queued requests) is called every
        ``progress_every`` wall seconds.

        A tick issues one captured step of every live job without waiting;
        the host synchronises every ``sync_every`` ticks (repairing capacity
        overflows) and on the tick that issues some job's last step, which
        is when that job is finished."""
        t0 = time.perf_counter()
        stats = ServeStats()
        tick_clock = clock == "ticks"
        if tick_clock:
            clock = lambda: float(stats.ticks)  # noqa: E731
        clock = clock or (lambda: time.perf_counter() - t0)
        pending = sorted(requests, key=lambda r: r.arrival)
        queue: List[FitRequest] = []
        last_report = t0
        stalled_rounds = 0
        while pending or queue or self.live:
            now = clock()
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            for req in list(queue):  # arrival order; a later request that fits may pass one that does not
                if self._admit(req, now):
                    queue.remove(req)
            if progress is not None and time.perf_counter() - last_report >= progress_every:
                last_report = time.perf_counter()
                progress(stats, len(self.live), len(queue))
            if not self.live:
                if pending and not queue:
                    if tick_clock:  # nothing to run: advance the tick clock to the next arrival
                        stats.ticks = int(math.ceil(pending[0].arrival))
                        continue
                    wait = pending[0].arrival - clock()
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                    continue
                if queue and not pending and not self.live:
                    raise RuntimeError(f"request {queue[0].rid} does not fit the HBM budget")
                continue
            stats.peak_bytes = max(stats.peak_bytes, self.used_bytes())
            stats.peak_live = max(stats.peak_live, len(self.live))
            running = [l for l in self.live if not l.stalled]
            if self.max_active is not None or self.active_ms is not None:  # the earliest arrivals first
                running = sorted(running, key=lambda l: (l.req.arrival, l.req.rid))
                if self.max_active is not None:
                    running = running[:self.max_active]
                if self.active_ms is not None:
                    total, k = 0.0, 0
                    while k < len(running) and total < self.active_ms:
                        total += self.estimate_ms(running[k].req)
                        k += 1
                    running = running[:k]
            for l in running:   # one iteration of every live job that can run
                l.fit.step(1)
                l.issued += 1
            stats.ticks += 1
            stats.job_steps += len(running)
            finishing = any(l.issued >= l.req.steps for l in running)
            if not finishing and stats.ticks % self.sync_every and not any(l.stalled for l in self.live):
                continue
            for l in self.live:
                self._sync(l)
            torch.cuda.synchronize(self.device)
            # every live job waiting for memory: give the released cache a few
            # retries before giving up
            stalled_rounds = stalled_rounds + 1 if self.live and all(l.stalled for l in self.live) else 0
            if stalled_rounds > 3:
                raise RuntimeError("every live job is waiting for device memory to grow its intersection capacity")
            now = clock()
            still = []
            for l in self.live:
                if l.issued >= l.req.steps and not l.stalled:
                    loss = float(l.fit.loss.item())
                    stats.finished.append((l.req.rid, l.req.arrival, l.admitted, now, loss, l.req.beta))
                    if keep_results:
                        stats.results[l.req.rid] = l.fit.params.clone()
                else:
                    still.append(l)
            self.live = still
            if max_wall is not None and time.perf_counter() - t0 > max_wall:
                break
        stats.wall = time.perf_counter() - t0
        return stats


__all__ = ["ExplicitFit", "FitRequest", "FitServer", "ServeStats", "build_job", "job_config", "poisson_requests"]

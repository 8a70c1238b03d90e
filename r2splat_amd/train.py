"""Graph-captured fits of float32 Gaussian or Beta scenes to target images.

``GaussianFit`` owns one reconstruction job: a flat parameter buffer (means,
quats, scales, densities[, betas] as views), its gradient buffer, Adam state
and every intermediate of the fused pipeline, all allocated once. One step is

    fused prepare -> render -> MSE adjoint -> raster VJP -> parameter VJP
    [-> Adam]

issued on one stream with no host synchronisation, so it can be captured into
a hipGraph (``torch.cuda.CUDAGraph``) and replayed. The only data-dependent
size, the intersection count, is bounded by a capacity: a step whose count
exceeds it raises the overflow flag, renders nothing, and its Adam update is
skipped on the device (counted in ``adam_state[1]``). ``sync()`` reads the
flags, grows the capacity, re-captures and replays the skipped steps, so the
sequence of applied updates is exactly that of an unbounded run.

The reference has no training loop (r2splat/explicit.py leaves activations,
losses and optimizers to the caller); this is the batched-fit layer on top of
the same operators.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import torch

from . import fused
from ._beta import beta_path_log_grad, beta_path_ratio
from .hip import _backend as be
from .hip._backend import call, ptr

TILE = fused.TILE


class Workspace:
    """Intermediates of the fused pipeline (projected records, sort scratch,
    packed intersection records, intersection partials, parameter-VJP
    contributions). Jobs issued on the same stream run one after another and
    can share one workspace, sized for the largest of them; ``version``
    changes whenever a buffer is reallocated (captured graphs then hold stale
    pointers and must be re-captured)."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.version = 0
        self.scratch = self.packed = self.isect_grads = self.param_ws = None

    @staticmethod
    def _grow(t, nbytes, device):
        if t is not None and t.numel() >= nbytes:
            return t, False
        return torch.empty((max(1, int(nbytes)),), dtype=torch.uint8, device=device), True

    def ensure(self, C, N, W, H, tile_size, capacity, beta: int = 0):
        """Grows the buffers to the given sizes. All new buffers are allocated
        before any is replaced, so an allocation failure leaves the workspace
        (and graphs captured on it) as it was."""
        lib = be.lib()
        grown = {}
        for name, nbytes in (("scratch", lib.r2s_fused_scratch_bytes(C, N, W, H, tile_size, capacity, beta)),
                             ("packed", lib.r2s_fused_packed_bytes(C, N, capacity, beta)),
                             ("isect_grads", capacity * 8 * 4),
                             ("param_ws", lib.r2s_fused_param_workspace_bytes(C, N))):
            t, new = self._grow(getattr(self, name), nbytes, self.device)
            if new:
                grown[name] = t
        for name, t in grown.items():
            setattr(self, name, t)
        if grown:
            self.version += 1

    def nbytes(self) -> int:
        return sum(t.numel() for t in (self.scratch, self.packed, self.isect_grads, self.param_ws) if t is not None)


@dataclass
class AdamConfig:
    lr: float = 1e-2
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    # after each update, clamp scales to >= min_scale, densities to >= 0 and
    # betas to >= min_beta (the explicit API's domain); None leaves the raw update
    min_scale: Optional[float] = None
    min_beta: float = 1e-3


class GaussianFit:
    """One job: N Gaussians (or Beta primitives, when ``beta`` [N] is given),
    C views of H x W, MSE against ``target`` [C, H, W]. ``fwd_waves`` /
    ``bwd_waves`` None take the waves of the job's shape
    (``fused.config_for_shape``), as the explicit API does, so that a default
    fit steps exactly as the explicit path computes; 0 is the kernels' default
    (8)."""

    def __init__(self, means, quats, scales, densities, viewmats, Ks, target, *, eps: float = 1e-8,
                 radius_clip: Optional[float] = None, adam: Optional[AdamConfig] = None, capacity: Optional[int] = None,
                 headroom: float = 1.15, stream: Optional[torch.cuda.Stream] = None, tile_size: int = TILE,
                 fwd_waves: Optional[int] = None, bwd_waves: Optional[int] = None,
                 workspace: Optional[Workspace] = None, beta=None):
        device = means.device
        if means.dtype != torch.float32:
            raise TypeError("GaussianFit runs the float32 fused pipeline")
        self.device = device
        self.N = int(means.shape[0])
        self.C = int(viewmats.shape[0])
        self.H, self.W = int(target.shape[-2]), int(target.shape[-1])
        if tuple(target.shape) != (self.C, self.H, self.W):
            raise ValueError("target must be [C, H, W]")
        if radius_clip is None:  # the explicit API's default: 3 sigma (Gaussian), the unit support (Beta)
            radius_clip = 3.0 if beta is None else 1.0
        self.eps, self.radius_clip = float(eps), float(radius_clip)
        self.adam = adam
        self.headroom = headroom
        _, shape_fwd, shape_bwd = fused.config_for_shape(self.N, self.C, self.H, self.W, beta is not None)
        self.tile_size = int(tile_size)
        self.fwd_waves = shape_fwd if fwd_waves is None else int(fwd_waves)
        self.bwd_waves = shape_bwd if bwd_waves is None else int(bwd_waves)
        self.stream = stream
        N = self.N
        self.kind = int(beta is not None)
        self.params = torch.empty(((11 + self.kind) * N,), dtype=torch.float32, device=device)
        self.grads = torch.zeros_like(self.params)
        self.means, self.quats, self.scales, self.densities, self.beta = self._views(self.params)
        self.v_means, self.v_quats, self.v_scales, self.v_densities, self.v_beta = self._views(self.grads)
        for dst, src in zip((self.means, self.quats, self.scales, self.densities, self.beta),
                            (means, quats, scales, densities, beta)):
            if dst is not None:
                dst.copy_(src)
        self.beta_ratio = self.beta_log_grad = None
        if self.kind:
            self.beta_ratio = torch.empty((N,), dtype=torch.float32, device=device)
            self.beta_log_grad = torch.empty((N,), dtype=torch.float32, device=device)
        self.viewmats = viewmats.contiguous().to(device, torch.float32)
        self.Ks = Ks.contiguous().to(device, torch.float32)
        self.target = target.contiguous().to(device, torch.float32)
        if adam is not None:
            self.exp_avg = torch.zeros_like(self.params)
            self.exp_avg_sq = torch.zeros_like(self.params)
        self.adam_state = torch.zeros((2,), dtype=torch.int32, device=device)
        n_pix = self.C * self.H * self.W
        self.image = torch.empty((self.C, self.H, self.W), dtype=torch.float32, device=device)
        self.v_image = torch.empty_like(self.image)
        self.loss_partials = torch.empty((be.lib().r2s_mse_partials(n_pix),), dtype=torch.float32, device=device)
        self.loss = torch.zeros((1,), dtype=torch.float32, device=device)
        ntx, nty = -(-self.W // self.tile_size), -(-self.H // self.tile_size)
        self.tile_offsets = torch.empty((self.C, nty, ntx), dtype=torch.int32, device=device)
        self.tile_order = torch.empty((self.C * nty * ntx,), dtype=torch.int32, device=device)
        self.record_offsets = torch.empty((self.C * N + 1 + N,), dtype=torch.int32, device=device)
        self.counters = torch.zeros((2,), dtype=torch.int32, device=device)
        self.ws = workspace if workspace is not None else Workspace(device)
        self.graph = None
        self._graph_version = -1
        self._regrown = False
        self.steps_issued = 0
        self._alloc(int(capacity) if capacity else self._measure_capacity())

    # -- buffers --------------------------------------------------------------
    def _views(self, flat):
        N = self.N
        return (flat[: 3 * N].view(N, 3), flat[3 * N: 7 * N].view(N, 4), flat[7 * N: 10 * N].view(N, 3),
                flat[10 * N: 11 * N], flat[11 * N:] if self.kind else None)

    def _alloc(self, capacity: int):
        # the buffers grow first and the capacity is adopted only once they all
        # exist: an allocation failure leaves the previous (consistent) capacity
        capacity = max(1024, int(capacity))
        self.ws.ensure(self.C, self.N, self.W, self.H, self.tile_size, capacity, self.kind)
        self.capacity = capacity

    @property
    def graph_stale(self) -> bool:
        return self.graph is None or self._graph_version != self.ws.version

    def _measure_capacity(self) -> int:
        fwd = fused.prepare(self.means, self.quats, self.scales, self.densities, self.viewmats, self.Ks, self.W,
                            self.H, self.eps, self.radius_clip, tile_size=self.tile_size, beta=self.beta)
        return int(math.ceil(fwd.n_isects * self.headroom)) + 1024

    def resident_bytes(self) -> int:
        """Device bytes this job holds between steps (parameters, optimizer
        state, images, per-view offsets), excluding its workspace."""
        tensors = [self.params, self.grads, self.image, self.v_image, self.loss_partials, self.tile_offsets,
                   self.tile_order, self.record_offsets, self.target, self.counters, self.adam_state]
        if self.adam is not None:
            tensors += [self.exp_avg, self.exp_avg_sq]
        if self.kind:
            tensors += [self.beta_ratio, self.beta_log_grad]
        return sum(t.numel() * t.element_size() for t in tensors)

    def memory_bytes(self) -> int:
        """Resident bytes plus its (possibly shared) workspace."""
        return self.resident_bytes() + self.ws.nbytes()

    # -- one step -------------------------------------------------------------
    # The step in stages (each enqueued on the current stream): prepare, render,
    # loss, raster VJP, update (parameter VJP, Adam, domain clamp). A batch of
    # jobs may run the render and the raster VJP of many jobs as one packed
    # launch (``raster_job``) between the other stages of each job.
    def _issue_prepare(self):
        s = be.stream_of(self.device)
        ws = self.ws
        if self.kind:  # line-integral factors of the current betas (float64 math, as the explicit path)
            self.beta_ratio.copy_(beta_path_ratio(self.beta))
            self.beta_log_grad.copy_(beta_path_log_grad(self.beta))
        call("r2s_fused_prepare", ptr(self.means), ptr(self.quats), ptr(self.scales), ptr(self.densities),
             ptr(self.beta), ptr(self.beta_ratio), None, ptr(self.viewmats), ptr(self.Ks), self.C, self.N, self.W,
             self.H, self.tile_size, self.eps, self.radius_clip, self.capacity,
             ptr(ws.scratch), ws.scratch.numel(), ptr(ws.packed), ptr(self.tile_offsets), ptr(self.tile_order),
             ptr(self.record_offsets), ptr(self.counters), s)

    def _issue_render(self):
        call("r2s_fused_render", ptr(self.ws.packed), ptr(self.tile_offsets), ptr(self.tile_order),
             ptr(self.counters), self.capacity, self.C, self.W, self.H, self.tile_size, self.fwd_waves, self.kind,
             ptr(self.image), None, be.stream_of(self.device))

    def _issue_loss(self, with_loss: bool = True):
        call("r2s_mse_adjoint", ptr(self.image), ptr(self.target), self.image.numel(), ptr(self.v_image),
             ptr(self.loss_partials) if with_loss else None, ptr(self.loss) if with_loss else None,
             be.stream_of(self.device))

    def _issue_raster_vjp(self):
        call("r2s_fused_raster_vjp", ptr(self.ws.packed), ptr(self.tile_offsets), ptr(self.tile_order),
             ptr(self.counters), self.capacity, self.C, self.W, self.H, self.tile_size, self.bwd_waves, self.kind,
             ptr(self.v_image), None, 0, ptr(self.ws.isect_grads), be.stream_of(self.device))

    def raster_job(self) -> dict:
        """This job's entry of a packed render / raster-VJP table (r2s_raster_job fields)."""
        return dict(packed=ptr(self.ws.packed), tile_offsets=ptr(self.tile_offsets), tile_order=ptr(self.tile_order),
                    counters=ptr(self.counters), image=ptr(self.image), v_image=ptr(self.v_image),
                    isect_grads=ptr(self.ws.isect_grads), capacity=self.capacity, C=self.C, width=self.W,
                    height=self.H)

    def _issue(self, with_loss: bool = True):
        """Enqueues one step on the current stream (no host synchronisation)."""
        self._issue_prepare()
        self._issue_render()
        self._issue_loss(with_loss)
        self._issue_raster_vjp()
        self._issue_update()

    def _issue_update(self):
        s = be.stream_of(self.device)
        C, N, W, H = self.C, self.N, self.W, self.H
        ws = self.ws
        call("r2s_fused_param_vjp", ptr(self.means), ptr(self.quats), ptr(self.scales), ptr(self.densities),
             ptr(self.beta), ptr(self.beta_ratio), ptr(self.beta_log_grad), ptr(self.viewmats), ptr(self.Ks), C, N,
             W, H, self.eps, ptr(self.record_offsets), ptr(self.counters), self.capacity, ptr(ws.isect_grads),
             ptr(ws.param_ws), ws.param_ws.numel(), ptr(self.v_means), ptr(self.v_quats), ptr(self.v_scales),
             ptr(self.v_densities), ptr(self.v_beta), None, s)
        if self.adam is not None:
            a = self.adam
            call("r2s_adam_step", ptr(self.params), ptr(self.grads), ptr(self.exp_avg), ptr(self.exp_avg_sq),
                 self.params.numel(), a.lr, a.beta1, a.beta2, a.eps, ptr(self.adam_state), ptr(self.counters[1:]),
                 s)
            if a.min_scale is not None:
                call("r2s_clamp_min", ptr(self.scales), self.scales.numel(), float(a.min_scale), s)
                call("r2s_clamp_min", ptr(self.densities), self.densities.numel(), 0.0, s)
                if self.kind:
                    call("r2s_clamp_min", ptr(self.beta), self.beta.numel(), float(a.min_beta), s)

    def _ctx(self):
        return torch.cuda.stream(self.stream) if self.stream is not None else torch.cuda.device(self.device)

    def capture(self):
        """Captures one step into a hipGraph (after one eager warm-up step)."""
        with self._ctx():
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    self._issue()
            torch.cuda.current_stream(self.device).wait_stream(side)
        self.graph = g
        self._graph_version = self.ws.version
        return g

    def step(self, n: int = 1, graph: bool = True):
        """Issues n steps (graph replays when captured) without synchronising."""
        with self._ctx():
            for _ in range(n):
                if graph:
                    if self.graph_stale:
                        raise RuntimeError("capture() first (the workspace changed since the last capture)")
                    self.graph.replay()
                else:
                    self._issue()
        self.steps_issued += n

    def pending_skips(self) -> int:
        """Steps whose forward overflowed the capacity (after a synchronisation)."""
        count, overflow = (int(v) for v in self.counters.tolist())
        applied, skipped = (int(v) for v in self.adam_state.tolist())
        if self.adam is None:
            return 1 if overflow else 0
        return skipped

    def regrow(self):
        """Grows the capacity past the last step's intersection count (and past
        the current capacity, so repeated overflows always make progress); an
        allocation failure leaves the previous capacity and buffers. The skip
        counter is left alone: it is cleared only when the skipped steps are
        re-issued (``replay_skipped``)."""
        count = int(self.counters[0].item())
        self._alloc(int(math.ceil(max(count, self.capacity) * self.headroom)) + 1024)

    def replay_skipped(self, skipped: int, graph: bool):
        """Clears the skip counter and re-issues `skipped` steps; call it only
        once every buffer (and, with `graph`, the graph) they need exists."""
        self.adam_state[1] = 0
        for _ in range(skipped):
            self.step(1, graph=graph)

    def sync(self) -> int:
        """Waits for the issued steps; on an overflow, grows the capacity,
        re-captures and re-runs the skipped steps. Returns the steps applied.

        Each stage is all or nothing: when growing the buffers or re-capturing
        raises (out of device memory), the skip counter still holds the
        skipped steps and a later sync retries from there; a graph left stale
        by a failed capture (or by a shared workspace that grew) is
        re-captured before anything is replayed."""
        with self._ctx():
            torch.cuda.current_stream(self.device).synchronize()
            while True:
                skipped = self.pending_skips()
                had_graph = self.graph is not None
                if skipped == 0:
                    if had_graph and self.graph_stale:
                        self.capture()
                    return int(self.adam_state[0].item()) if self.adam is not None else self.steps_issued
                if not self._regrown:  # grow once per run of skipped steps, even if a capture fails after it
                    self.regrow()
                    self._regrown = True
                if had_graph and self.graph_stale:
                    self.capture()
                self._regrown = False
                self.replay_skipped(skipped, graph=had_graph)
                torch.cuda.current_stream(self.device).synchronize()

    def loss_value(self) -> float:
        return float(self.loss.item())


__all__ = ["AdamConfig", "GaussianFit", "Workspace"]

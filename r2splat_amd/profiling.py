"""Per-stage GPU timing with HIP events recorded on the launch stream.

Disabled by default (zero cost beyond a flag test). When enabled, each
``stage(name)`` block records a start/end event pair on the current stream of
the active device — the stream every operator launches on — so the measured
duration is that of the kernels issued inside the block.
"""

from __future__ import annotations

import contextlib
from collections import defaultdict

import torch

_enabled = False
_pending: list = []
_totals = defaultdict(float)
_counts = defaultdict(int)


def enable(flag: bool = True) -> None:
    global _enabled
    _enabled = flag


def reset() -> None:
    _pending.clear()
    _totals.clear()
    _counts.clear()


@contextlib.contextmanager
def stage(name: str):
    if not _enabled:
        yield
        return
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    try:
        yield
    finally:
        end.record()
        _pending.append((name, start, end))


def collect() -> dict:
    """Synchronises pending events; returns {stage: (mean_ms, count)}."""
    for name, start, end in _pending:
        end.synchronize()
        _totals[name] += start.elapsed_time(end)
        _counts[name] += 1
    _pending.clear()
    return {k: (_totals[k] / _counts[k], _counts[k]) for k in _totals}

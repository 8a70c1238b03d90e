"""ctypes front-end of the C/OpenMP CPU restatement — TEST INFRASTRUCTURE ONLY.

Used by tests (checked against the float64 oracle) and by bench.py's
cpu_baseline leg (timed on the host cores).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import LIB, build

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        I = ctypes.c_int
        F = ctypes.c_float
        lib.r2s_cpu_iteration.argtypes = [P, P, P, P, P, P, I, I, I, I, I, F, F, P, P, P, I]
        lib.r2s_cpu_iteration.restype = ctypes.c_int64
        lib.r2s_cpu_max_threads.restype = I
        _lib = lib
    return _lib


def max_threads() -> int:
    return int(_load().r2s_cpu_max_threads())


def iteration(means, quats, scales, densities, viewmats, Ks, height, width, target, *, tile=16, eps=1e-8,
              clip=3.0, threads=0):
    """One fwd + MSE + bwd iteration on the CPU. numpy float32 inputs.
    Returns (image [C,H,W], grads dict, n_isects)."""
    lib = _load()
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (means, quats, scales, densities, viewmats, Ks, target)]
    N, C = arrs[0].shape[0], arrs[4].shape[0]
    image = np.zeros((C, height, width), dtype=np.float32)
    grads = np.zeros((N, 11), dtype=np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    n_is = lib.r2s_cpu_iteration(p(arrs[0]), p(arrs[1]), p(arrs[2]), p(arrs[3]), p(arrs[4]), p(arrs[5]), N, C,
                                 height, width, tile, eps, clip, p(arrs[6]), p(image), p(grads), threads)
    g = {"means": grads[:, 0:3], "quats": grads[:, 3:7], "scales": grads[:, 7:10], "densities": grads[:, 10]}
    return image, g, int(n_is)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1

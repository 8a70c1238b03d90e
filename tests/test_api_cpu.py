"""Host-side contracts that need no GPU: public surface, validation errors, and
the C ABI library (loads, exports every declared symbol, matches the ctypes table)."""

import ctypes
import math
import re
from pathlib import Path

import pytest
import torch

import r2splat_amd
from r2splat_amd import rasterize_fwd, voxelize_fwd
from r2splat_amd._validation import options, positive_ints
from r2splat_amd.hip import _backend

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "r2splat_hip.h"


def test_public_api_is_small():
    assert r2splat_amd.__all__ == ["__version__", "rasterize_fwd", "rasterize_bwd", "voxelize_fwd",
                                   "voxelize_bwd", "ParameterGradients"]
    assert not hasattr(r2splat_amd, "rasterize")


@pytest.mark.parametrize("shape", [(0, 3), (True, 3), (2.5, 3), (3,), (2**30, 4)])
def test_invalid_output_shape_rejected(shape):
    with pytest.raises(ValueError, match="image_size"):
        positive_ints("image_size", shape, 2)


@pytest.mark.parametrize("eps,radius,detach", [(-1, None, False), (math.nan, None, False), (0, 0, False),
                                               (0, math.inf, False), (0, 1, "yes")])
def test_invalid_numeric_options_rejected(eps, radius, detach):
    with pytest.raises((ValueError, TypeError)):
        options(None, eps, radius, detach)


def test_radius_clip_defaults():
    assert options(None, 0.0, None, False) == (0.0, 3.0)
    assert options(torch.ones(1), 0.0, None, False) == (0.0, 1.0)


def test_cpu_inputs_fail_before_native_dispatch():
    values = (torch.zeros(1, 3), torch.ones(1, 4), torch.ones(1, 3), torch.ones(1))
    with pytest.raises(ValueError, match="means must be a CUDA tensor"):
        voxelize_fwd(*values, (3, 4, 5), volume_extent=(1, 1, 1))
    with pytest.raises(ValueError, match="means must be a CUDA tensor"):
        rasterize_fwd(*values, torch.eye(4)[None], torch.eye(3)[None], (4, 4))


@pytest.mark.parametrize("dtype", [torch.float16, torch.int32])
def test_unsupported_precision_is_a_type_error(dtype):
    with pytest.raises(TypeError, match="dtype"):
        voxelize_fwd(torch.zeros(1, 3, dtype=dtype), None, None, None, (3, 4, 5), volume_extent=(1, 1, 1))


def _header_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(r2s_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_backend.library_path()))
    symbols = _header_symbols()
    assert len(symbols) >= 30
    missing = [s for s in symbols if not hasattr(lib, s)]
    assert missing == []


def test_ctypes_table_covers_the_header():
    assert sorted(_backend.SIGNATURES) == _header_symbols()


def test_library_version_and_workspace_queries_need_no_gpu():
    lib = _backend.lib(require_gpu=False)
    assert lib.r2s_abi_version() == 5
    assert lib.r2s_scan_workspace_bytes(0) > 0
    big = lib.r2s_sort_workspace_bytes(10**7, 15)
    assert big >= 4 * 4 * 10**7


def test_fused_size_queries_need_no_gpu():
    """The packed buffer: `capacity` 8-byte list entries (rounded up to 256
    bytes), then C*N records of 32 bytes (Gaussian), 48 bytes (Gaussian with
    uncertainty) or 96 bytes (Beta: the 48-byte record and its double-precision
    form); the scratch grows with the intersection capacity only on the sort
    path (tile grids too large for direct placement)."""
    lib = _backend.lib(require_gpu=False)
    assert [lib.r2s_fused_record_bytes(k) for k in range(4)] == [32, 96, 48, 96]
    for C, N, cap, kind in ((4, 1000, 1000, 0), (64, 100_000, 30_000_000, 0), (3, 7, 5, 3)):
        entries = -(-cap * 8 // 256) * 256
        records = -(-C * N * lib.r2s_fused_record_bytes(kind) // 256) * 256
        assert lib.r2s_fused_packed_bytes(C, N, cap, kind) == entries + records
    # 256 tiles of 16 pixels per camera: direct placement, no capacity-sized sort buffers
    small = lib.r2s_fused_scratch_bytes(64, 100_000, 256, 256, 16, 10**6, 0)
    assert lib.r2s_fused_scratch_bytes(64, 100_000, 256, 256, 16, 10**8, 0) == small
    # 1024 tiles of 16 pixels per camera: sort + pack, keys and values per intersection
    a = lib.r2s_fused_scratch_bytes(4, 1000, 512, 512, 16, 10**6, 0)
    b = lib.r2s_fused_scratch_bytes(4, 1000, 512, 512, 16, 2 * 10**6, 0)
    assert b - a >= 8 * 10**6


def test_workspace_growth_is_all_or_nothing(monkeypatch):
    """A capacity growth that runs out of memory part way leaves every buffer
    and the version (which graphs captured on the workspace check) as they
    were; the next growth succeeds and bumps the version once."""
    from r2splat_amd import train

    load = _backend.lib
    monkeypatch.setattr(train.be, "lib", lambda: load(require_gpu=False))  # size queries only
    ws = train.Workspace("cpu")
    ws.ensure(4, 1000, 64, 64, 16, 4096)
    names = ("scratch", "packed", "isect_grads", "param_ws")
    before = {n: getattr(ws, n) for n in names}
    version = ws.version
    real_empty, calls = torch.empty, []

    def empty_failing_second(*args, **kwargs):
        calls.append(args)
        if len(calls) == 2:
            raise torch.OutOfMemoryError("simulated")
        return real_empty(*args, **kwargs)

    monkeypatch.setattr(torch, "empty", empty_failing_second)
    with pytest.raises(torch.OutOfMemoryError):
        ws.ensure(4, 1000, 64, 64, 16, 1 << 20)  # grows the list entries and the partials
    monkeypatch.setattr(torch, "empty", real_empty)
    assert len(calls) == 2 and ws.version == version
    assert all(getattr(ws, n) is before[n] for n in names)
    ws.ensure(4, 1000, 64, 64, 16, 1 << 20)
    assert ws.version == version + 1
    assert ws.isect_grads.numel() >= (1 << 20) * 32 and ws.packed.numel() >= (1 << 20) * 8
    assert ws.scratch is before["scratch"] and ws.param_ws is before["param_ws"]  # sizes independent of capacity


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_operators_raise_without_gpu():
    from r2splat_amd import hip

    assert not hip.is_cuda_extension_available()
    with pytest.raises(RuntimeError):
        hip.parameter_transform_fwd(torch.ones(1, 4), torch.ones(1, 3))


def test_bench_refuses_more_than_one_gpu():
    """The benchmark measures one GPU; asking for more exits with status 2 and a
    JSON message, before anything touches a device."""
    import json
    import subprocess
    import sys

    proc = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)
    assert proc.returncode == 2
    assert "one GPU" in json.loads(proc.stdout.strip().splitlines()[-1])["error"]


def test_shape_configurations_are_valid_kernel_configurations():
    """fused.config_for_shape gives a tile edge the fused kernels support and
    wave counts they are built for, for every shape of the serving mix and
    beyond; Beta primitives and images up to 256^2 keep 16-pixel tiles."""
    from r2splat_amd import fused
    from r2splat_amd.batched import KernelConfig, heuristic_config

    for n in (1, 1000, 10_000, 60_000, 400_000, 2_000_000):
        for views in (1, 16, 32, 64, 128, 256):
            for side in (32, 128, 256, 300, 512, 1024):
                for beta in (False, True):
                    tile, fw, bw = fused.config_for_shape(n, views, side, side, beta)
                    assert tile in (8, 16, 32) and fw in (4, 8, 16) and bw in (4, 8, 16)
                    if beta or side < 512:
                        assert tile == 16
                    assert heuristic_config(n, views, side, side, beta) == KernelConfig(tile, fw, bw)


def test_python_volume_frames_are_checked_on_the_host():
    """Volume-frame vectors given as Python numbers: rounded to the
    primitives' precision before the finiteness and sign checks, cached per
    value (signed zeros apart), other forms left to the tensor checks."""
    from r2splat_amd import explicit

    like = torch.zeros(1, 3)
    a = explicit._constant_vector((0.0, -0.0, 2), like)
    b = explicit._constant_vector((0.0, 0.0, 2), like)
    assert a is not b and torch.equal(a[0], b[0]) and a[0][1].item() == 0.0
    assert torch.signbit(a[0][1]) and not torch.signbit(b[0][1])
    assert explicit._constant_vector((0.0, -0.0, 2), like) is a
    assert a[2:] == (False, True)  # finite; not all positive
    assert explicit._constant_vector((1.0, 2.0, 3.0), like)[2:] == (False, False)
    assert explicit._constant_vector((float("nan"), 1.0, 1.0), like)[2:] == (True, False)
    assert explicit._constant_vector((1e39, 1.0, 1.0), like)[2] is True  # overflows float32
    assert explicit._constant_vector((1e39, 1.0, 1.0), like.double())[2] is False
    assert explicit._constant_vector((1e39, 1.0, 1.0), like.double())[1].dtype == torch.float64
    for other in ((1.0, 1.0), [1.0, 1.0, 1.0, 1.0], (10**400, 1, 1), (True, 1, 1), "abc", torch.ones(3)):
        assert explicit._constant_vector(other, like) is None

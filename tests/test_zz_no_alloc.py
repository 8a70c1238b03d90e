"""The allocation-free MSE adjoint (r2s_mse_adjoint_noalloc) against the
workspace entry point (r2s_mse_adjoint with caller partials), through the C
ABI: two calls in a row, the second smaller, each bitwise equal; sizes beyond
the static partial buffer are rejected with a message."""

import pytest
import torch

from r2splat_amd.hip import _backend

pytestmark = pytest.mark.gpu

LIMIT = (1 << 16) * 2048  # static partials x pixels per partial


def _workspace_call(lib, img, tgt, stream):
    n = img.numel()
    v = torch.empty_like(img)
    loss = torch.empty(1, dtype=torch.float32, device=img.device)
    parts = torch.empty(lib.r2s_mse_partials(n), dtype=torch.float32, device=img.device)
    assert lib.r2s_mse_adjoint(img.data_ptr(), tgt.data_ptr(), n, v.data_ptr(), parts.data_ptr(), loss.data_ptr(),
                               stream) == 0
    return v, loss


def _noalloc_call(lib, img, tgt, stream):
    n = img.numel()
    v = torch.empty_like(img)
    loss = torch.full((1,), float("nan"), dtype=torch.float32, device=img.device)
    rc = lib.r2s_mse_adjoint_noalloc(img.data_ptr(), tgt.data_ptr(), n, v.data_ptr(), loss.data_ptr(), stream)
    assert rc == 0, lib.r2s_last_error().decode()
    return v, loss


def test_noalloc_mse_adjoint_matches_workspace_entry_point(device):
    lib = _backend.lib()
    stream = torch.cuda.current_stream(device).cuda_stream
    g = torch.Generator(device=device).manual_seed(7)
    # a large call, then a smaller one: the static partials keep the first
    # call's contents, which the second must not pick up
    for n in (64 * 256 * 256 + 12_345, 70_001):
        print(f"noalloc mse: n={n}", flush=True)
        img = torch.rand(n, generator=g, device=device)
        tgt = torch.rand(n, generator=g, device=device)
        v_ref, loss_ref = _workspace_call(lib, img, tgt, stream)
        v, loss = _noalloc_call(lib, img, tgt, stream)
        torch.cuda.synchronize(device)
        assert torch.equal(v, v_ref)
        assert torch.equal(loss, loss_ref)
        expected = float(((img.double() - tgt.double()) ** 2).mean())
        assert abs(float(loss) - expected) <= 1e-5 * expected


def test_noalloc_mse_adjoint_rejects_sizes_beyond_its_buffer(device):
    lib = _backend.lib()
    stream = torch.cuda.current_stream(device).cuda_stream
    x = torch.zeros(4, device=device)
    rc = lib.r2s_mse_adjoint_noalloc(x.data_ptr(), x.data_ptr(), LIMIT + 1, x.data_ptr(), x.data_ptr(), stream)
    assert rc != 0
    assert "static partial buffer" in lib.r2s_last_error().decode()
    rc = lib.r2s_mse_adjoint_noalloc(x.data_ptr(), x.data_ptr(), 4, x.data_ptr(), None, stream)
    assert rc != 0
    assert "loss is required" in lib.r2s_last_error().decode()
    torch.cuda.synchronize(device)


def test_noalloc_mse_adjoint_empty_input_zeroes_the_loss(device):
    lib = _backend.lib()
    stream = torch.cuda.current_stream(device).cuda_stream
    loss = torch.full((1,), 5.0, device=device)
    assert lib.r2s_mse_adjoint_noalloc(None, None, 0, None, loss.data_ptr(), stream) == 0
    torch.cuda.synchronize(device)
    assert float(loss) == 0.0

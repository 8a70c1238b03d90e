"""Input value checks of the explicit operators: the native batch
(r2s_value_checks) against the torch form of each predicate, element by
element kind and precision (NaN, infinities, signed zeros, all-zero rows, the
last element of a multi-workgroup buffer, empty buffers, more than one launch
of 16); and the operators' ValueError messages with their precedence (the
reference's order: required fields' finiteness, their domains, the optional
fields, then the cameras or the volume frame), for volume frames given as
tensors or as Python numbers (checked on the host, rounded to the primitives'
precision first)."""

import ctypes

import pytest
import torch

from r2splat_amd import _validation as check
from r2splat_amd.explicit import rasterize_fwd, voxelize_fwd
from r2splat_amd.hip import _backend as be

pytestmark = pytest.mark.gpu

KINDS = (be.CHECK_NONFINITE, be.CHECK_NONPOSITIVE, be.CHECK_NEGATIVE, be.CHECK_ZERO_ROW)


def _buffers(device, dtype):
    """(name, tensor) pairs, each with at most one special element."""
    big = 3_000_017  # several workgroups of the scan, odd length
    out = [("clean", torch.rand(1000, device=device, dtype=dtype) + 0.5),
           ("empty", torch.empty(0, device=device, dtype=dtype))]
    for name, v in (("nan", float("nan")), ("inf", float("inf")), ("-inf", float("-inf")), ("zero", 0.0),
                    ("-zero", -0.0), ("negative", -1e-30), ("tiny", 1e-45)):
        for at in (0, 517, 999):
            t = torch.rand(1000, device=device, dtype=dtype) + 0.5
            t[at] = v
            out.append((f"{name}@{at}", t))
    t = torch.rand(big, device=device, dtype=dtype) + 0.5
    t[-1] = float("nan")
    out.append(("big nan at the end", t))
    t = torch.rand(big, device=device, dtype=dtype) + 0.5
    t[big // 2] = -2.0
    out.append(("big negative", t))
    for rows, zero_row in ((250, None), (250, 0), (250, 249), (750_000, 749_999)):
        t = torch.rand(rows * 4, device=device, dtype=dtype) + 0.5
        t.view(rows, 4)[:, 1:3] = 0.0  # partly zero rows are not zero rows
        if zero_row is not None:
            t.view(rows, 4)[zero_row] = torch.tensor([0.0, -0.0, 0.0, 0.0], dtype=dtype)
        out.append((f"rows zero at {zero_row}", t))
    return out


def _native_flags(entries, device):
    """entries: (tensor, kind, row); one r2s_value_checks call per 16."""
    flags = torch.full((len(entries),), 7, dtype=torch.int32, device=device)
    for lo in range(0, len(entries), be.MAX_VALUE_CHECKS):
        chunk = entries[lo: lo + be.MAX_VALUE_CHECKS]
        table = (be.ValueCheck * len(chunk))()
        for e, (t, kind, row) in zip(table, chunk):
            e.data, e.n, e.dtype, e.kind, e.row = t.data_ptr(), t.numel(), be.dtype_code(t.dtype), kind, row
        be.call("r2s_value_checks", ctypes.addressof(table), len(chunk), flags[lo:].data_ptr(),
                be.stream_of(device))
    return flags.tolist()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_native_checks_equal_torch_predicates(device, dtype):
    entries, names = [], []
    for name, t in _buffers(device, dtype):
        for kind in KINDS:
            entries.append((t, kind, 4 if kind == be.CHECK_ZERO_ROW and t.numel() % 4 == 0 else 1))
            names.append((name, kind))
    got = _native_flags(entries, device)
    want = [int(bool(check._element_predicate(t, kind, row))) for t, kind, row in entries]
    assert got == want, [n for n, g, w in zip(names, got, want) if g != w]
    assert 0 < sum(want) < len(want)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_native_camera_checks_equal_the_matrix_predicates(device, dtype):
    """Intrinsics [C, 3, 3] and viewmats [C, 4, 4] with one entry changed per
    case (NaN, infinities, signed zeros, a tiny skew), against the predicates
    written on the matrices' entries."""
    C = 40
    K0 = torch.tensor([[20.0, 0, 8], [0, 30, 9], [0, 0, 1]], dtype=dtype).repeat(C, 1, 1)
    V0 = torch.eye(4, dtype=dtype).repeat(C, 1, 1)
    V0[:, :3, 3] = torch.rand(C, 3, dtype=dtype)
    cases = [("K", K0.clone()), ("V", V0.clone())]
    for (r, c) in ((0, 0), (1, 1), (0, 1), (1, 0), (2, 0), (2, 1), (2, 2), (0, 2)):
        for v in (0.0, -0.0, -1.0, 1e-30, 1.0, float("nan"), float("inf")):
            K = K0.clone()
            K[C - 1 if (r + c) % 2 else 0, r, c] = v
            cases.append(("K", K))
    for c in range(4):
        for v in (0.0, -0.0, 1.0, 1e-30, float("nan")):
            V = V0.clone()
            V[17, 3, c] = v
            cases.append(("V", V))
    entries, want = [], []
    for kind_of, m in cases:
        m = m.to(device)
        if kind_of == "K":
            entries += [(m, be.CHECK_K_FOCAL, 9), (m, be.CHECK_K_PINHOLE, 9)]
            want += [bool((m[:, (0, 1), (0, 1)] <= 0).any()),
                     bool((m[:, (0, 1, 2, 2), (1, 0, 0, 1)] != 0).any() | (m[:, 2, 2] != 1).any())]
        else:
            entries.append((m, be.CHECK_VIEW_ROW, 16))
            want.append(bool((m[:, 3] != m.new_tensor([0, 0, 0, 1])).any()))
    got = _native_flags(entries, device)
    assert got == [int(w) for w in want]
    assert [int(bool(check._element_predicate(t, k, r))) for t, k, r in entries] == got
    assert 0 < sum(want) < len(want)


def test_native_checks_reject_bad_tables(device):
    t = torch.ones(8, device=device)
    table = (be.ValueCheck * 1)()
    table[0].data, table[0].n, table[0].dtype, table[0].kind, table[0].row = t.data_ptr(), 8, 0, be.CHECK_ZERO_ROW, 3
    flags = torch.zeros(1, dtype=torch.int32, device=device)
    with pytest.raises(RuntimeError, match="row must be positive and divide n"):
        be.call("r2s_value_checks", ctypes.addressof(table), 1, flags.data_ptr(), be.stream_of(device))
    table[0].kind, table[0].row = be.CHECK_K_FOCAL, 4
    with pytest.raises(RuntimeError, match="intrinsics rows hold 9 elements"):
        be.call("r2s_value_checks", ctypes.addressof(table), 1, flags.data_ptr(), be.stream_of(device))
    table[0].kind, table[0].row = be.CHECK_VIEW_ROW, 8
    with pytest.raises(RuntimeError, match="viewmat rows hold 16 elements"):
        be.call("r2s_value_checks", ctypes.addressof(table), 1, flags.data_ptr(), be.stream_of(device))
    table[0].kind = 9
    with pytest.raises(RuntimeError, match="unknown kind"):
        be.call("r2s_value_checks", ctypes.addressof(table), 1, flags.data_ptr(), be.stream_of(device))
    many = (be.ValueCheck * 17)()
    with pytest.raises(RuntimeError, match="at most 16"):
        be.call("r2s_value_checks", ctypes.addressof(many), 17, flags.data_ptr(), be.stream_of(device))


def test_batch_reports_the_first_failing_check_in_order(device):
    clean = torch.ones(100, device=device)
    bad = clean.clone()
    bad[3] = float("nan")
    checks = check.ValueChecks()
    for k in range(20):  # two native launches
        checks.finite(f"c{k}", bad if k in (17, 19) else clean)
    checks.known("host check", True)
    with pytest.raises(ValueError, match="^c17 must contain only finite values$"):
        checks.run()
    checks = check.ValueChecks()
    checks.finite("a", clean)
    checks.known("host check", True)
    checks.add("torch check", lambda: torch.tensor(True, device=device))
    with pytest.raises(ValueError, match="^host check$"):
        checks.run()
    checks = check.ValueChecks()
    checks.finite("a", clean)
    checks.finite("other precision", clean.double())
    checks.add("torch check", lambda: torch.tensor(True, device=device))
    checks.finite("b", bad)
    with pytest.raises(ValueError, match="^torch check$"):
        checks.run()
    checks = check.ValueChecks()
    checks.element("zero row", clean.view(25, 4), be.CHECK_ZERO_ROW, 4)
    checks.element("negative", clean.double(), be.CHECK_NEGATIVE)
    checks.run()  # nothing bad: no error


def _prims(device, dtype, beta=True, unc=True):
    g = torch.Generator(device="cpu").manual_seed(5)
    n = 64
    return {
        "means": (torch.rand(n, 3, generator=g, dtype=dtype) - 0.5).to(device),
        "quats": torch.randn(n, 4, generator=g, dtype=dtype).to(device),
        "scales": (torch.rand(n, 3, generator=g, dtype=dtype) * 0.05 + 0.01).to(device),
        "densities": torch.rand(n, generator=g, dtype=dtype).to(device),
        "beta": (torch.rand(n, generator=g, dtype=dtype) + 0.5).to(device) if beta else None,
        "uncertainty": (torch.rand(n, generator=g, dtype=dtype) * 0.1).to(device) if unc else None,
    }


def _voxelize(p, extent=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0)):
    return voxelize_fwd(p["means"], p["quats"], p["scales"], p["densities"], (8, 9, 10), volume_extent=extent,
                        volume_center=center, beta=p["beta"], uncertainty=p["uncertainty"])


def _rasterize(p, device, dtype, fx=20.0):
    viewmats = torch.eye(4, device=device, dtype=dtype)[None].clone()
    viewmats[0, 2, 3] = 2.0
    Ks = torch.tensor([[[fx, 0, 8], [0, 20, 8], [0, 0, 1]]], device=device, dtype=dtype)
    return rasterize_fwd(p["means"], p["quats"], p["scales"], p["densities"], viewmats, Ks, (16, 16),
                         beta=p["beta"], uncertainty=p["uncertainty"])


BREAKS = [
    ("means", lambda t: t[5].fill_(float("nan")), "means must contain only finite values"),
    ("quats", lambda t: t[7].zero_(), "quats must have nonzero norm"),
    ("quats", lambda t: t[7, 2].fill_(float("inf")), "quats must contain only finite values"),
    ("scales", lambda t: t[63, 1].fill_(0.0), "scales must be strictly positive physical lengths"),
    ("densities", lambda t: t[0].fill_(-1e-3), "densities must be nonnegative"),
    ("beta", lambda t: t[9].fill_(0.0), "beta must contain positive 3D exponents"),
    ("uncertainty", lambda t: t[9].fill_(-0.5), "uncertainty must be nonnegative relative variance"),
    ("uncertainty", lambda t: t[9].fill_(float("nan")), "uncertainty must contain only finite values"),
]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("field,breaks,message", BREAKS)
def test_operators_report_each_bad_field(device, dtype, field, breaks, message):
    for op in ("voxelize", "rasterize"):
        p = _prims(device, dtype)
        breaks(p[field])
        with pytest.raises(ValueError, match=f"^{message}$"):
            _voxelize(p) if op == "voxelize" else _rasterize(p, device, dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_volume_frame_checks_and_precedence(device, dtype):
    p = _prims(device, dtype)
    _voxelize(p)  # valid
    for extent in ((1.0, 0.0, 1.0), (1, 1, -2)):
        with pytest.raises(ValueError, match="^volume_extent must contain positive physical lengths$"):
            _voxelize(p, extent=extent)
        with pytest.raises(ValueError, match="^volume_extent must contain positive physical lengths$"):
            _voxelize(p, extent=torch.tensor(extent, device=device, dtype=dtype))
    with pytest.raises(ValueError, match="^volume_extent must contain only finite values$"):
        _voxelize(p, extent=(1.0, float("inf"), 1.0))
    with pytest.raises(ValueError, match="^volume_center must contain only finite values$"):
        _voxelize(p, center=(0.0, float("nan"), 0.0))
    with pytest.raises(ValueError, match="^volume_center must contain only finite values$"):
        _voxelize(p, center=torch.tensor([0.0, float("nan"), 0.0], device=device, dtype=dtype))
    # a finite Python number that overflows the primitives' precision is not finite there
    if dtype == torch.float32:
        with pytest.raises(ValueError, match="^volume_extent must contain only finite values$"):
            _voxelize(p, extent=(1e39, 1.0, 1.0))
    # the primitives come first, the extent's finiteness before its sign, and a
    # tensor centre's check keeps its place between the two
    bad = dict(p)
    bad["densities"] = p["densities"].clone()
    bad["densities"][3] = -1.0
    with pytest.raises(ValueError, match="^densities must be nonnegative$"):
        _voxelize(bad, extent=(float("nan"), -1.0, 1.0))
    with pytest.raises(ValueError, match="^volume_extent must contain only finite values$"):
        _voxelize(p, extent=(float("nan"), -1.0, 1.0))
    with pytest.raises(ValueError, match="^volume_center must contain only finite values$"):
        _voxelize(p, extent=(1.0, -1.0, 1.0),
                  center=torch.tensor([float("inf"), 0.0, 0.0], device=device, dtype=dtype))
    with pytest.raises(ValueError, match="^volume_extent must contain positive physical lengths$"):
        _voxelize(p, extent=torch.tensor([1.0, -1.0, 1.0], device=device, dtype=dtype), center=(0.0, 0.0, 0.0))
    # the same Python vector in the other precision is a separate cached frame
    other = torch.float64 if dtype == torch.float32 else torch.float32
    _voxelize(_prims(device, other), extent=(1.0, 1.0, 1.0))
    _voxelize(p, extent=(1.0, 1.0, 1.0))


def test_camera_checks_still_run(device):
    p = _prims(device, torch.float32)
    _rasterize(p, device, torch.float32)
    with pytest.raises(ValueError, match="^Ks must have positive fx and fy$"):
        _rasterize(p, device, torch.float32, fx=0.0)
    bad = dict(p)
    bad["means"] = p["means"].clone()
    bad["means"][0, 0] = float("nan")
    with pytest.raises(ValueError, match="^means must contain only finite values$"):
        _rasterize(bad, device, torch.float32, fx=0.0)
    with pytest.raises(ValueError, match="^Ks must contain only finite values$"):
        _rasterize(p, device, torch.float32, fx=float("nan"))

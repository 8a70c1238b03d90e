"""Pins the CPU oracle (and the host-side Beta factor) before it is trusted.

Each check is chosen so that a plausible slip (a dropped term, a swapped
index, a wrong sign) fails: golden tables from the reference's tests, closed
forms, independent formulations and brute force on tiny inputs.
"""

import json
import math
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import splat as O
from r2splat_amd._beta import beta_path_log_grad, beta_path_ratio

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(autouse=True)
def _float64_default():
    previous = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(previous)


# --------------------------------------------------------------------------- Beta
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_beta_factor_matches_160_digit_table(dtype):
    rows = json.load(open(GOLDEN / "beta_path_factor.json"))["rows"]
    data = torch.tensor(rows, dtype=dtype)
    gamma = data[:, 0].clone().requires_grad_()
    ratio = beta_path_ratio(gamma)
    auto = torch.autograd.grad(ratio.log().sum(), gamma)[0]
    tol = 3e-6 if dtype == torch.float32 else 2e-12
    assert ratio.dtype == dtype
    torch.testing.assert_close(ratio, data[:, 1], atol=0, rtol=tol)
    torch.testing.assert_close(auto, data[:, 2], atol=0, rtol=tol)
    torch.testing.assert_close(beta_path_log_grad(gamma.detach()), data[:, 2], atol=0, rtol=tol)


def test_beta_factor_huge_exponent_stays_finite():
    gamma = torch.tensor([1e100], requires_grad=True)
    ratio = beta_path_ratio(gamma)
    grad = torch.autograd.grad(ratio.log().sum(), gamma)[0]
    torch.testing.assert_close(ratio, torch.tensor([1.0 / math.sqrt(2e100)]), atol=0, rtol=1e-13)
    torch.testing.assert_close(grad, torch.tensor([-5e-101]), atol=0, rtol=1e-13)


def test_beta_factor_integer_exponents_closed_form():
    # integral of (1 - t^2)^n over [-1, 1] = 2 prod_k 2k / (2k + 1)
    n = [1, 2, 8]
    expected = torch.tensor([2 * math.prod(2 * k / (2 * k + 1) for k in range(1, m + 1)) / math.sqrt(2 * math.pi)
                             for m in n])
    torch.testing.assert_close(beta_path_ratio(torch.tensor(n, dtype=torch.float64)), expected, atol=0, rtol=1e-13)
    torch.testing.assert_close(O.beta_ratio(torch.tensor(n, dtype=torch.float64)), expected, atol=0, rtol=1e-13)


def test_oracle_beta_ratio_matches_table_for_moderate_exponents():
    rows = [r for r in json.load(open(GOLDEN / "beta_path_factor.json"))["rows"] if r[0] <= 40]
    data = torch.tensor(rows)
    torch.testing.assert_close(O.beta_ratio(data[:, 0]), data[:, 1], atol=0, rtol=1e-12)


def test_beta_log_grad_is_derivative_of_log_ratio():
    gamma = torch.tensor([0.8, 3.5, 9.0, 19.99, 20.01, 300.0])
    h = 1e-6
    fd = (beta_path_ratio(gamma + h).log() - beta_path_ratio(gamma - h).log()) / (2 * h)
    torch.testing.assert_close(beta_path_log_grad(gamma), fd, atol=1e-9, rtol=1e-7)


# --------------------------------------------------------------------------- covariance
def test_covariance_matches_cross_product_form_and_is_scale_invariant():
    g = torch.Generator().manual_seed(0)
    q = torch.randn(64, 4, generator=g)
    s = 0.1 + torch.rand(64, 3, generator=g)
    a = O.covariance(q, s)
    torch.testing.assert_close(a, O.covariance_cross_product(q, s), atol=1e-13, rtol=1e-12)
    torch.testing.assert_close(a, O.covariance(-3.7 * q, s), atol=1e-13, rtol=1e-12)
    torch.testing.assert_close(a, a.transpose(-1, -2))
    # eigenvalues are the squared scales
    ev = torch.linalg.eigvalsh(a)
    torch.testing.assert_close(ev, torch.sort(s.square(), dim=-1).values, atol=1e-12, rtol=1e-10)


def test_pack_roundtrip():
    m = O.covariance(torch.randn(5, 4), torch.rand(5, 3) + 0.1)
    torch.testing.assert_close(O.unpack_sym(O.pack_sym(m)), m)


# --------------------------------------------------------------------------- projection
def _axis_scene(scales, eps=0.0, fx=12.0, fy=10.0, cx=4.5, cy=3.5, z=2.0):
    means = torch.tensor([[0.0, 0.0, z]])
    quats = torch.tensor([[1.0, 0.0, 0.0, 0.0]])
    view = torch.eye(4)[None]
    K = torch.tensor([[[fx, 0, cx], [0, fy, cy], [0, 0, 1.0]]])
    return means, quats, torch.tensor([scales]), view, K


def test_projection_closed_form_on_axis():
    sx, sy, sz = 0.3, 0.45, 0.6
    means, quats, scales, view, K = _axis_scene([sx, sy, sz])
    m2, c2, vis = O.project(means, O.covariance(quats, scales), view, K, 9, 7)
    assert bool(vis.all())
    torch.testing.assert_close(m2[0, 0], torch.tensor([4.5, 3.5]))
    expected = torch.diag(torch.tensor([(12 * sx / 2) ** 2, (10 * sy / 2) ** 2, sz ** 2]))
    torch.testing.assert_close(c2[0, 0], expected, atol=1e-14, rtol=1e-13)
    for eps in (0.0, 1e-3):
        conics, mu = O.conics_mu(c2, eps)
        torch.testing.assert_close(conics[0, 0], torch.tensor([1 / (expected[0, 0] + eps), 0.0,
                                                                1 / (expected[1, 1] + eps)]))
        torch.testing.assert_close(mu[0, 0], torch.tensor(math.sqrt(2 * math.pi * (sz ** 2 + eps))))


def test_projection_jacobian_rows_are_pixel_map_derivative_and_ray_direction():
    # Off-axis, unclamped point: rows 0-1 of J are d(pixel)/d(camera point), row 2 the unit ray.
    g = torch.Generator().manual_seed(3)
    R = torch.linalg.qr(torch.randn(3, 3, generator=g))[0]
    if torch.det(R) < 0:
        R[:, 0] = -R[:, 0]
    view = torch.eye(4)
    view[:3, :3] = R
    view[:3, 3] = torch.tensor([0.1, -0.2, 3.0])
    K = torch.tensor([[[30.0, 0, 16], [0, 25.0, 12], [0, 0, 1]]])
    mean = torch.tensor([[0.2, 0.1, -0.3]])
    cov = O.covariance(torch.randn(1, 4, generator=g), torch.rand(1, 3, generator=g) + 0.1)
    m2, c2, _ = O.project(mean, cov, view[None], K, 32, 24)
    pc = (R @ mean[0] + view[:3, 3]).detach()

    def pixel(p):
        return torch.stack([30 * p[0] / p[2] + 16, 25 * p[1] / p[2] + 12])

    J2 = torch.autograd.functional.jacobian(pixel, pc)
    ray = pc / pc.norm()
    J = torch.cat([J2, ray[None]], 0)
    Sc = R @ cov[0] @ R.T
    torch.testing.assert_close(c2[0, 0], J @ Sc @ J.T, atol=1e-12, rtol=1e-11)
    torch.testing.assert_close(m2[0, 0], pixel(pc))


def test_fov_clamp_limits_ray_point():
    # A point far outside the field of view: tx is clamped to z * (1.15 W - cx) / fx.
    means = torch.tensor([[50.0, 0.0, 2.0]])
    cov = O.covariance(torch.tensor([[1.0, 0, 0, 0]]), torch.tensor([[0.1, 0.1, 0.1]]))
    view = torch.eye(4)[None]
    K = torch.tensor([[[12.0, 0, 4.5], [0, 10.0, 3.5], [0, 0, 1]]])
    _, c2, _ = O.project(means, cov, view, K, 9, 7)
    tx = 2.0 * (1.15 * 9 - 4.5) / 12.0
    J0 = torch.tensor([12 / 2.0, 0.0, -12 * tx / 4.0])
    torch.testing.assert_close(c2[0, 0, 0, 0], 0.01 * (J0 @ J0), rtol=1e-12, atol=0)


def test_conics_mu_matches_determinant_ratio():
    g = torch.Generator().manual_seed(1)
    A = torch.randn(16, 3, 3, generator=g)
    S = A @ A.transpose(-1, -2) + 0.1 * torch.eye(3)
    for eps in (0.0, 0.3):
        conics, mu = O.conics_mu(S, eps)
        M = S + eps * torch.eye(3)
        P = torch.linalg.inv(M[:, :2, :2])
        torch.testing.assert_close(conics, P[:, [0, 0, 1], [0, 1, 1]])
        torch.testing.assert_close(mu, torch.sqrt(2 * math.pi * torch.det(M) / torch.det(M[:, :2, :2])))


def test_conics_mu_degenerate_is_zero():
    S = torch.zeros(2, 3, 3)
    S[1] = torch.ones(3, 3)
    conics, mu = O.conics_mu(S, 0.0)
    assert float(conics.abs().sum()) == 0.0 and float(mu.abs().sum()) == 0.0


# --------------------------------------------------------------------------- line integrals
@pytest.mark.parametrize("beta", [None, 3.5])
def test_centred_line_integral_closed_form(beta):
    sz, dens = 0.4, 1.3
    means, quats, scales, view, K = _axis_scene([0.2, 0.3, sz], cx=0.5, cy=0.5, fx=1.0, fy=1.0)
    b = None if beta is None else torch.tensor([beta])
    img = O.rasterize_scene(means, quats, scales, torch.tensor([dens]), view, K, (1, 1), beta=b, eps=0.0)
    expected = dens * math.sqrt(2 * math.pi) * sz
    if beta is not None:
        expected *= math.exp(math.lgamma(0.5) + math.lgamma(beta + 1) - math.lgamma(beta + 1.5)) / math.sqrt(2 * math.pi)
    torch.testing.assert_close(img.reshape(()), torch.tensor(expected), atol=0, rtol=1e-13)


def test_gaussian_projection_matches_numerical_ray_integral():
    # Orthographic limit check: a small primitive far away integrates along z to
    # density * sqrt(2 pi) * sqrt(Schur complement of the camera covariance).
    g = torch.Generator().manual_seed(5)
    q = torch.randn(1, 4, generator=g)
    s = torch.tensor([[0.05, 0.08, 0.11]])
    cov = O.covariance(q, s)[0]
    z = torch.linspace(-1, 1, 200001)
    prec = torch.linalg.inv(cov)
    # density at (0, 0, z) relative to the centre
    dens = torch.exp(-0.5 * prec[2, 2] * z ** 2)
    integral = torch.trapezoid(dens, z)
    schur = cov[2, 2] - cov[2, :2] @ torch.linalg.solve(cov[:2, :2], cov[:2, 2])
    assert abs(integral.item() - math.sqrt(2 * math.pi / prec[2, 2].item())) < 1e-9
    assert abs(1 / prec[2, 2].item() - schur.item()) < 1e-12


def test_dense_rasterize_matches_python_loops():
    g = torch.Generator().manual_seed(2)
    C, N, H, W = 2, 3, 5, 6
    m2 = torch.rand(C, N, 2, generator=g) * torch.tensor([W, H])
    A = torch.randn(C, N, 2, 2, generator=g)
    P = A @ A.transpose(-1, -2) * 0.2 + 0.1 * torch.eye(2)
    conics = P[..., [0, 0, 1], [0, 1, 1]]
    opac = torch.rand(C, N, generator=g)
    mus = torch.rand(C, N, generator=g)
    beta2d = 1 + torch.rand(C, N, generator=g)
    unc = torch.rand(N, generator=g)
    for beta in (None, beta2d):
        img, var = O.rasterize(m2, conics, opac, mus, W, H, beta2d=beta, unc=unc)
        for c in range(C):
            for i in range(H):
                for j in range(W):
                    tot = tv = 0.0
                    for n in range(N):
                        d = torch.tensor([j + 0.5, i + 0.5]) - m2[c, n]
                        Pm = P[c, n]
                        if beta is None:
                            s = 0.5 * float(d @ Pm @ d)
                            a = float(opac[c, n] * mus[c, n]) * math.exp(-s)
                            ok = s >= 0 and abs(a) >= 1e-5
                        else:
                            d2 = float(d @ Pm @ d)
                            ok = 0 <= d2 < 1
                            a = float(opac[c, n] * mus[c, n]) * (1 - d2) ** float(beta[c, n]) if ok else 0.0
                            ok = ok and abs(a) >= 1e-5
                        if ok:
                            tot += a
                            tv += float(unc[n]) * a * a
                    assert abs(img[c, i, j].item() - tot) < 1e-13
                    assert abs(var[c, i, j].item() - tv) < 1e-13


# --------------------------------------------------------------------------- binning
def test_tile_bounds_3d_known_answers():
    kat = json.load(open(GOLDEN / "tile_bounds_3d.json"))
    nx, ny, nz = kat["n_tiles_xyz"]
    for case in kat["cases"]:
        for dt in (np.float32, np.float64):
            means = np.array([case["mean"]], dtype=dt)
            radii = np.full((1, 3), case["radius"], dtype=np.int32)
            counts, ids, flat = O.isect_tiles_3d(means, radii, kat["tile_size"], nx, ny, nz)
            assert counts.tolist() == [len(case["expected_ids"])], case["id"]
            assert ids.tolist() == case["expected_ids"], case["id"]
            assert flat.tolist() == [0] * len(case["expected_ids"])
            off = O.offsets(ids, nx * ny * nz)
            expect = np.searchsorted(np.array(case["expected_ids"], dtype=np.int64), np.arange(24))
            assert off.tolist() == expect.tolist()


def test_tile_binning_2d_matches_brute_force():
    rng = np.random.default_rng(7)
    C, N, T, W, H = 3, 40, 4, 22, 17
    ntx, nty = -(-W // T), -(-H // T)
    for dt in (np.float32, np.float64):
        m2 = (rng.uniform(-6, [W + 6, H + 6], size=(C, N, 2))).astype(dt)
        radii = rng.integers(0, 9, size=(C, N, 2)).astype(np.int32)
        counts, ids, flat = O.isect_tiles_2d(m2, radii, T, ntx, nty)
        recs = []
        for c in range(C):
            for n in range(N):
                rx, ry = int(radii[c, n, 0]), int(radii[c, n, 1])
                if rx <= 0 or ry <= 0:
                    continue
                cx, cy = m2[c, n, 0] / dt(T), m2[c, n, 1] / dt(T)
                rxt, ryt = dt(rx) / dt(T), dt(ry) / dt(T)

                def cl(v, hi):
                    v = float(v)
                    return 0 if (not math.isfinite(v) or v <= 0) else (hi if v >= hi else int(v))

                x0, x1 = cl(np.floor(cx - rxt), ntx), cl(np.ceil(cx + rxt), ntx)
                y0, y1 = cl(np.floor(cy - ryt), nty), cl(np.ceil(cy + ryt), nty)
                cnt = 0
                for y in range(y0, y1):
                    for x in range(x0, x1):
                        recs.append((c * ntx * nty + y * ntx + x, c * N + n))
                        cnt += 1
                assert counts[c, n] == cnt
        recs.sort()  # (key, value) lexicographic == stable sort of values emitted in order
        assert ids.tolist() == [r[0] for r in recs]
        assert flat.tolist() == [r[1] for r in recs]


def test_tile_mask_matches_rects():
    rng = np.random.default_rng(8)
    m2 = rng.uniform(0, 40, size=(2, 5, 2)).astype(np.float32)
    radii = rng.integers(1, 12, size=(2, 5, 2)).astype(np.int32)
    mask = O.tile_mask_2d(m2, radii, 16, 40, 33)
    x0, x1, y0, y1 = O.tile_rects_2d(m2, radii, 16, 3, 3)
    for c in range(2):
        for n in range(5):
            for i in range(33):
                for j in range(40):
                    inside = x0[c, n] <= j // 16 < x1[c, n] and y0[c, n] <= i // 16 < y1[c, n]
                    assert mask[c, n, i, j] == inside


# --------------------------------------------------------------------------- voxels / TV
def test_voxel_sample_positions_and_centre_value():
    means = torch.tensor([[0.1, -0.2, 2.0]])
    quats = torch.tensor([[1.0, 0, 0, 0]])
    scales = torch.tensor([[0.001, 0.001, 0.001]])
    # single voxel centred on the primitive
    vol = O.voxelize(means, quats, scales, torch.tensor([2.0]), (1, 1, 1), (1.0, 1.0, 1.0), (0.1, -0.2, 2.0), eps=0.0)
    assert abs(vol.item() - 2.0) < 1e-14
    # 3 voxels along x: centres at center_x + ((k + 1/2)/3 - 1/2) * extent_x
    scales = torch.tensor([[0.5, 0.5, 0.5]])
    vol = O.voxelize(means, quats, scales, torch.tensor([1.0]), (1, 1, 3), (3.0, 1.0, 1.0), (0.1, -0.2, 2.0), eps=0.0)
    expected = torch.exp(-0.5 * torch.tensor([-1.0, 0.0, 1.0]) ** 2 / 0.25)
    torch.testing.assert_close(vol[0, 0], expected, atol=0, rtol=1e-13)


def test_volume_mapping_precision_matches_world_frame():
    means = torch.tensor([[0.1, -0.1, 0.7]])
    cov = O.covariance(torch.tensor([[0.9, 0.1, -0.2, 0.3]]), torch.tensor([[0.2, 0.3, 0.4]]))
    ext, cen, shape = torch.tensor([2.0, 3.0, 4.0]), torch.tensor([0.2, -0.3, 1.0]), (5, 9, 13)
    mapped, P, ok = O.volume_mapping(means, cov, ext, cen, shape, 1e-3)
    r = torch.tensor([13.0, 9.0, 5.0]) / ext
    torch.testing.assert_close(mapped, (means - cen) * r + torch.tensor([13.0, 9.0, 5.0]) / 2)
    world_prec = torch.linalg.inv(cov[0] + torch.diag(1e-3 / r ** 2))
    torch.testing.assert_close(P[0], world_prec / r[:, None] / r[None, :], atol=1e-12, rtol=1e-10)
    assert bool(ok.all())


def test_volume_tv_matches_loops():
    g = torch.Generator().manual_seed(4)
    v = torch.randn(2, 3, 4, 5, generator=g)
    tot, edges = 0.0, 0
    for b in range(2):
        for d in range(3):
            for h in range(4):
                for w in range(5):
                    for dd, dh, dw in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                        if d + dd < 3 and h + dh < 4 and w + dw < 5:
                            tot += abs(float(v[b, d + dd, h + dh, w + dw] - v[b, d, h, w]))
                            edges += 1
    assert abs(O.volume_tv(v).item() - tot / edges) < 1e-13


def test_volume_tv_iso_known_answers():
    """Isotropic TV: a constant volume gives sqrt(eps) and a zero gradient; a
    linear ramp a x + b y + c z gives sqrt(a^2 + b^2 + c^2 + eps) wherever all
    three forward differences exist; a single row reduces to the mean of
    sqrt(dx^2 + eps); the gradient matches central differences."""
    eps = 1e-8
    const = torch.full((1, 3, 4, 5), 2.5, dtype=torch.float64, requires_grad=True)
    tv = O.volume_tv_iso(const, eps)
    assert abs(tv.item() - eps ** 0.5) < 1e-15
    (g,) = torch.autograd.grad(tv, const)
    assert float(g.abs().max()) == 0.0
    d, h, w = torch.meshgrid(torch.arange(4.0), torch.arange(5.0), torch.arange(6.0), indexing="ij")
    a, b, c = 0.3, -1.2, 0.7
    ramp = (a * w + b * h + c * d).double()[None]
    t_interior = (a * a + b * b + c * c + eps) ** 0.5
    expect = 0.0
    for dd in range(4):
        for hh in range(5):
            for ww in range(6):
                sq = (c * c if dd < 3 else 0) + (b * b if hh < 4 else 0) + (a * a if ww < 5 else 0)
                expect += (sq + eps) ** 0.5
    assert abs(O.volume_tv_iso(ramp, eps).item() - expect / ramp.numel()) < 1e-14
    assert abs(expect / ramp.numel() - t_interior) > 1e-3  # the boundary terms matter
    row = torch.tensor([[[[0.0, 1.0, -2.0, 0.5]]]], dtype=torch.float64)
    dxs = [1.0, -3.0, 2.5, 0.0]
    assert abs(O.volume_tv_iso(row, eps).item() - sum((x * x + eps) ** 0.5 for x in dxs) / 4) < 1e-15
    gen = torch.Generator().manual_seed(9)
    v = torch.randn(2, 3, 4, 3, generator=gen, dtype=torch.float64, requires_grad=True)
    (g,) = torch.autograd.grad(O.volume_tv_iso(v, 1e-6), v)
    h_ = 1e-6
    for idx in [(0, 0, 0, 0), (1, 2, 3, 2), (0, 1, 2, 1), (1, 0, 3, 0)]:
        vp, vm = v.detach().clone(), v.detach().clone()
        vp[idx] += h_
        vm[idx] -= h_
        fd = (O.volume_tv_iso(vp, 1e-6) - O.volume_tv_iso(vm, 1e-6)).item() / (2 * h_)
        assert abs(fd - g[idx].item()) < 1e-7


# --------------------------------------------------------------------------- radii (cull rules)
def test_radii_closed_form_on_axis():
    """On the optical axis cov2d[0,0] = (fx sx / z)^2 and cov2d[1,1] = (fy sy / z)^2,
    so rx = ceil(clip sqrt((fx sx / z)^2 + eps)) (WorldToImageProjectionFwd.cu:250-276):
    checked either side of an integer so a dropped eps or a floor fails."""
    fx, fy, z, clip, eps = 12.0, 10.0, 2.0, 3.0, 0.25
    for sx, sy in ((0.3, 0.45), (0.1, 0.05), (0.8, 0.02)):
        means, quats, scales, view, K = _axis_scene([sx, sy, 0.5], fx=fx, fy=fy, z=z)
        _, c2, _ = O.project(means, O.covariance(quats, scales), view, K, 9, 7)
        packed = O.pack_sym(c2).numpy()
        radii = O.radii_2d(packed, np.array([[[4.5, 3.5]]]), eps, clip, 9, 7)
        rx = math.ceil(clip * math.sqrt((fx * sx / z) ** 2 + eps))
        ry = math.ceil(clip * math.sqrt((fy * sy / z) ** 2 + eps))
        assert radii[0, 0].tolist() == [rx, ry]
    # variance chosen so that clip * sqrt(var + eps) lands just below / above 5
    for var, expect in ((25.0 / 9.0 - eps - 1e-9, 5), (25.0 / 9.0 - eps + 1e-9, 6)):
        packed = np.array([[[var, 0.0, 0.0, 1.0, 0.0, 1.0]]])
        assert O.radii_2d(packed, np.array([[[4.5, 3.5]]]), eps, clip, 9, 7)[0, 0, 0] == expect


def test_radii_cull_rules():
    """Zero radii for a non-positive-definite 2x2 block, non-finite input and
    each of the four off-screen tests; the boundary itself (mean +- r exactly on
    the image edge) is culled (<= 0, >= width/height)."""
    W, H = 20, 10
    base = [4.0, 0.0, 0.0, 1.0, 0.0, 1.0]  # sigma_x 2, sigma_y 1: rx = ceil(3 sqrt(4)) = 6, ry = 3

    def radii(cov, mean, eps=0.0):
        return O.radii_2d(np.array([[cov]]), np.array([[mean]]), eps, 3.0, W, H)[0, 0].tolist()

    assert radii(base, [10.0, 5.0]) == [6, 3]
    assert radii([4.0, 2.0, 0.0, 1.0, 0.0, 1.0], [10.0, 5.0]) == [0, 0]  # det = 0
    assert radii([4.0, 2.5, 0.0, 1.0, 0.0, 1.0], [10.0, 5.0]) == [0, 0]  # det < 0
    assert radii([4.0, 2.0, 0.0, 1.0, 0.0, 1.0], [10.0, 5.0], eps=0.1)[0] > 0  # eps restores det > 0
    assert radii([math.nan, 0.0, 0.0, 1.0, 0.0, 1.0], [10.0, 5.0]) == [0, 0]
    assert radii([math.inf, 0.0, 0.0, 1.0, 0.0, 1.0], [10.0, 5.0]) == [0, 0]
    # left: mean_x + rx <= 0
    assert radii(base, [-6.0, 5.0]) == [0, 0] and radii(base, [-5.5, 5.0]) == [6, 3]
    # right: mean_x - rx >= W
    assert radii(base, [W + 6.0, 5.0]) == [0, 0] and radii(base, [W + 5.5, 5.0]) == [6, 3]
    # top: mean_y + ry <= 0
    assert radii(base, [10.0, -3.0]) == [0, 0] and radii(base, [10.0, -2.5]) == [6, 3]
    # bottom: mean_y - ry >= H
    assert radii(base, [10.0, H + 3.0]) == [0, 0] and radii(base, [10.0, H + 2.5]) == [6, 3]


def test_projection_culls_points_at_or_behind_the_near_plane():
    """z <= 0.01 gives zero means2d and cov2d and is not visible
    (WorldToImageProjectionFwd.cu:129-136)."""
    cov = O.covariance(torch.tensor([[1.0, 0, 0, 0]] * 3), torch.full((3, 3), 0.1))
    means = torch.tensor([[0.0, 0.0, 0.01], [0.0, 0.0, 0.0100001], [0.3, 0.0, -1.0]])
    K = torch.tensor([[[12.0, 0, 4.5], [0, 10.0, 3.5], [0, 0, 1]]])
    m2, c2, vis = O.project(means, cov, torch.eye(4)[None], K, 9, 7)
    assert vis[0].tolist() == [False, True, False]
    assert float(m2[0, 0].abs().sum()) == 0.0 and float(c2[0, 0].abs().sum()) == 0.0
    assert float(m2[0, 2].abs().sum()) == 0.0 and float(c2[0, 2].abs().sum()) == 0.0


# --------------------------------------------------------------------------- 3D tile mask
def test_tile_mask_3d_matches_loops():
    rng = np.random.default_rng(11)
    N, T, shape = 6, 4, (7, 9, 13)  # (D, H, W)
    means = rng.uniform(-3, [16, 12, 10], size=(N, 3))
    radii = rng.integers(0, 6, size=(N, 3)).astype(np.int32)
    radii[0] = [2, 0, 3]  # a zero radius on one axis: empty
    mask = O.tile_mask_3d(means, radii, T, shape)
    D, H, W = shape
    for n in range(N):
        spans = []
        for k, size in enumerate((W, H, D)):
            nt = -(-size // T)
            lo = math.floor(means[n, k] / T - radii[n, k] / T)
            hi = math.ceil(means[n, k] / T + radii[n, k] / T)
            spans.append((min(max(lo, 0), nt), min(max(hi, 0), nt)))
        live = bool((radii[n] > 0).all())
        for z in range(D):
            for y in range(H):
                for x in range(W):
                    inside = live and all(s[0] <= v // T < s[1] for s, v in zip(spans, (x, y, z)))
                    assert mask[n, z, y, x] == inside, (n, z, y, x)
    assert not mask[0].any()


# --------------------------------------------------------------------------- voxel Beta / variance
def _centred_pair():
    """The reference's two-primitive scene centred on a 1x1x1 volume
    (its tests/test_beta_api.py:14-19,26-28)."""
    means = torch.tensor([[0.0, 0, 2], [0.0, 0, 2]])
    quats = torch.tensor([[2.0, 0, 0, 0], [3.0, 0, 0, 0]])
    scales = torch.tensor([[0.2, 0.3, 0.4], [0.3, 0.4, 0.5]])
    dens = torch.tensor([1.2, 0.8])
    return means, quats, scales, dens


@pytest.mark.parametrize("detach", [False, True])
def test_voxelize_beta_variance_centred_closed_form(detach):
    """Every primitive sits on the voxel centre: d2 = 0, alpha = density, so
    mean = sum rho, variance = sum q rho^2; with adjoints (1, 0.4) the VJPs are
    d/drho = 1 + 0.8 q rho (1 when detached), d/dbeta = 0 (log 1), d/dq = 0.4 rho^2."""
    means, quats, scales, dens = _centred_pair()
    rho = dens.clone().requires_grad_()
    beta = torch.tensor([3.5, 1.25], requires_grad=True)
    q = torch.tensor([0.2, 0.3], requires_grad=True)
    mean, var = O.voxelize(means, quats, scales, rho, (1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 2.0), beta=beta,
                           uncertainty=q, eps=0.0, detach=detach)
    assert abs(mean.item() - 2.0) < 1e-15
    assert abs(var.item() - (0.2 * 1.44 + 0.3 * 0.64)) < 1e-15
    g_rho, g_beta, g_q = torch.autograd.grad(mean.sum() + 0.4 * var.sum(), (rho, beta, q))
    factor = 1.0 + (0.0 if detach else 0.8) * q.detach() * dens
    torch.testing.assert_close(g_rho, factor, atol=1e-15, rtol=0)
    assert float(g_beta.abs().max()) == 0.0
    torch.testing.assert_close(g_q, 0.4 * dens.square(), atol=1e-15, rtol=0)


def test_voxelize_beta_off_centre_axis_aligned_by_hand():
    """Axis-aligned ellipsoid (distinct semi-axes) on a non-cubic grid with
    distinct extents: d2 = sum_k (x_k - m_k)^2 / s_k^2 at voxel centres
    center + ((index + 1/2) / n - 1/2) extent, alpha = rho (1 - d2)^beta kept
    for 0 <= d2 < 1 and alpha >= 1e-5; variance q alpha^2. Catches swapped
    axes, a missing half-voxel offset and a Gaussian/Beta branch mix-up."""
    mean = (0.13, -0.21, 1.92)
    s = (0.7, 0.45, 0.3)
    rho, beta, q = 0.9, 2.5, 0.07
    shape = (3, 5, 4)  # (D, H, W)
    extent, center = (1.6, 1.1, 0.75), (0.05, -0.1, 2.0)
    vol, var = O.voxelize(torch.tensor([mean]), torch.tensor([[1.0, 0, 0, 0]]), torch.tensor([s]),
                          torch.tensor([rho]), shape, extent, center, beta=torch.tensor([beta]),
                          uncertainty=torch.tensor([q]), eps=0.0)
    D, H, W = shape
    kept = 0
    for z in range(D):
        for y in range(H):
            for x in range(W):
                p = [center[k] + ((i + 0.5) / n - 0.5) * extent[k] for k, (i, n) in enumerate(((x, W), (y, H), (z, D)))]
                d2 = sum((p[k] - mean[k]) ** 2 / s[k] ** 2 for k in range(3))
                a = rho * (1 - d2) ** beta if 0 <= d2 < 1 else 0.0
                a = a if a >= 1e-5 else 0.0
                kept += a > 0
                assert abs(vol[z, y, x].item() - a) < 1e-14, (z, y, x)
                assert abs(var[z, y, x].item() - q * a * a) < 1e-15
    assert 0 < kept < D * H * W  # both branches of the support test are exercised


def test_voxelize_cutoffs_1e6_gaussian_1e5_beta():
    """The thresholds themselves: a single voxel at distance d2 from the
    primitive is kept iff alpha >= 1e-6 (Gaussian, VoxelizeToVoxelsFwd.cu:213)
    or alpha >= 1e-5 (Beta, :575)."""
    s = 0.5

    def one(d2, beta=None):
        dx = s * math.sqrt(d2)  # offset along x giving that d2
        b = None if beta is None else torch.tensor([beta])
        return O.voxelize(torch.tensor([[dx, 0.0, 2.0]]), torch.tensor([[1.0, 0, 0, 0]]), torch.tensor([[s, s, s]]),
                          torch.tensor([1.0]), (1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 2.0), beta=b, eps=0.0).item()

    g_edge = -2 * math.log(1e-6)  # e^(-d2/2) = 1e-6
    assert one(g_edge * (1 - 1e-9)) > 0 and one(g_edge * (1 + 1e-9)) == 0.0
    beta = 2.0
    b_edge = 1 - 1e-5 ** (1 / beta)  # (1 - d2)^2 = 1e-5
    assert one(b_edge * (1 - 1e-9), beta) > 0 and one(b_edge * (1 + 1e-9), beta) == 0.0
    assert one(1.0, 0.5) == 0.0 and one(0.999999, 0.5) > 0  # support ends at d2 = 1


# --------------------------------------------------------------------------- absgrad
def _abs_scene(beta):
    g = torch.Generator().manual_seed(21)
    C, N, W, H = 2, 5, 9, 7
    m2 = torch.rand(C, N, 2, generator=g) * torch.tensor([W, H])
    A = torch.randn(C, N, 2, 2, generator=g)
    P = A @ A.transpose(-1, -2) * 0.1 + 0.05 * torch.eye(2)
    conics = P[..., [0, 0, 1], [0, 1, 1]]
    opac, mus = 0.3 + torch.rand(C, N, generator=g), 0.5 + torch.rand(C, N, generator=g)
    b2 = 1 + 2 * torch.rand(C, N, generator=g) if beta else None
    return m2, conics, opac, mus, b2, W, H


@pytest.mark.parametrize("beta", [False, True])
def test_abs_mean2d_grad_signed_parts_sum_to_the_gradient(beta):
    """The per-pixel partials behind absgrad, summed with their signs, are the
    gradient of sum(v * image) with respect to means2d (autograd of rasterize)."""
    m2, conics, opac, mus, b2, W, H = _abs_scene(beta)
    v = torch.linspace(-1.0, 1.5, 2 * H * W).reshape(2, H, W)
    leaf = m2.clone().requires_grad_()
    img = O.rasterize(leaf, conics, opac, mus, W, H, beta2d=b2)
    (grad,) = torch.autograd.grad((img * v).sum(), leaf)
    ab = O.abs_mean2d_grad(m2, conics, opac, mus, W, H, v, beta2d=b2)
    assert bool((ab >= grad.abs() - 1e-12).all())  # |sum| <= sum of |parts|
    assert float((ab - grad.abs()).max()) > 1e-3    # and strictly larger somewhere: signs differ across pixels
    # a positive adjoint and every kept pixel right of / below the mean: all parts share a sign
    one = torch.tensor([[[-0.5, -0.5]]])
    cone = torch.tensor([[[0.4, 0.0, 0.6]]])
    vp = torch.ones(1, H, W)
    leaf = one.clone().requires_grad_()
    bb = None if b2 is None else torch.tensor([[1.5]])
    img = O.rasterize(leaf, cone, torch.tensor([[1.0]]), torch.tensor([[1.0]]), W, H, beta2d=bb)
    if beta:  # the Beta support of conic 0.4 / 0.6 reaches ~1.6 px: pick a wider one
        cone = torch.tensor([[[0.05, 0.0, 0.08]]])
        img = O.rasterize(leaf, cone, torch.tensor([[1.0]]), torch.tensor([[1.0]]), W, H, beta2d=bb)
    (grad,) = torch.autograd.grad((img * vp).sum(), leaf)
    ab = O.abs_mean2d_grad(one, cone, torch.tensor([[1.0]]), torch.tensor([[1.0]]), W, H, vp, beta2d=bb)
    assert float(grad.abs().min()) > 0
    torch.testing.assert_close(ab, grad.abs(), atol=1e-14, rtol=1e-12)


def test_abs_mean2d_grad_variance_adjoint_enters_unless_detached():
    m2, conics, opac, mus, _, W, H = _abs_scene(False)
    v = torch.zeros(2, H, W)
    vv = torch.ones(2, H, W)
    unc = torch.full((5,), 0.5)
    with_var = O.abs_mean2d_grad(m2, conics, opac, mus, W, H, v, unc=unc, v_var=vv)
    detached = O.abs_mean2d_grad(m2, conics, opac, mus, W, H, v, unc=unc, v_var=vv, detach=True)
    assert float(detached.abs().max()) == 0.0 and float(with_var.min()) >= 0 and float(with_var.max()) > 0
    # with v = 0 the pixel adjoint is 2 q alpha v_var: the parts are those of the variance sum q alpha^2
    leaf = m2.clone().requires_grad_()
    _, var = O.rasterize(leaf, conics, opac, mus, W, H, unc=unc)
    (g,) = torch.autograd.grad(var.sum(), leaf)
    assert bool((with_var >= g.abs() - 1e-12).all())

"""Known answers on the GPU: closed forms that pin the drop-in API without the
oracle in the loop.

* The reference's centred Beta + uncertainty cases (its
  tests/test_beta_api.py:24-67): two primitives on the optical axis of a 1x1
  image / at the centre of a 1x1x1 volume; mean = sum rho path, variance =
  sum q (rho path)^2, and the density, beta and uncertainty VJPs in closed form.
* An off-centre axis-aligned Beta ellipsoid on a non-cubic voxel grid, each
  voxel computed by hand.
* The voxel cut-offs (1e-6 Gaussian, 1e-5 Beta) either side of the threshold.
* Projection radii: the on-axis closed form and the off-screen / near-plane culls.
"""

import math

import pytest
import torch

import r2splat_amd
from r2splat_amd import hip as ops

pytestmark = pytest.mark.gpu
DTYPES = (torch.float32, torch.float64)


def _pair(device, dtype):
    means = torch.tensor([[0, 0, 2], [0, 0, 2]], device=device, dtype=dtype)
    quats = means.new_tensor([[2, 0, 0, 0], [3, 0, 0, 0]])
    scales = means.new_tensor([[0.2, 0.3, 0.4], [0.3, 0.4, 0.5]])
    densities = means.new_tensor([1.2, 0.8])
    return means, quats, scales, densities


def _beta_path(b):
    return math.exp(math.lgamma(0.5) + math.lgamma(b + 1) - math.lgamma(b + 1.5))


@pytest.mark.parametrize("operation", ["rasterize", "voxelize"])
@pytest.mark.parametrize("with_uncertainty", [False, True])
@pytest.mark.parametrize("dtype", DTYPES)
def test_centred_beta_and_uncertainty_closed_forms(device, operation, with_uncertainty, dtype):
    values = _pair(device, dtype)
    beta = values[0].new_tensor([3.5, 99, 1.25, 99])[::2]  # a strided view: the API copies it
    q = values[0].new_tensor([0.2, 0.3]) if with_uncertainty else None
    print(f"known answer: {operation} unc={with_uncertainty} {dtype}", flush=True)
    if operation == "rasterize":
        view = torch.eye(4, device=device, dtype=dtype).repeat(2, 1, 1)
        K = values[0].new_tensor([[[1, 0, 0.5], [0, 1, 0.5], [0, 0, 1]]]).repeat(2, 1, 1)
        out, ctx = r2splat_amd.rasterize_fwd(*values, view, K, (1, 1), beta=beta, uncertainty=q, eps=0)
        cameras = 2
        # mu = scale_z B(1/2, beta + 1): the line integral of (1 - d^2)^beta through the centre
        path = values[2][:, 2] * values[0].new_tensor([_beta_path(3.5), _beta_path(1.25)])
        dlog = (torch.special.digamma(beta.double() + 1) - torch.special.digamma(beta.double() + 1.5)).to(dtype)
    else:
        out, ctx = r2splat_amd.voxelize_fwd(*values, (1, 1, 1), volume_extent=(1, 1, 1), volume_center=(0, 0, 2),
                                            beta=beta, uncertainty=q, eps=0)
        cameras = 1
        path = torch.ones_like(beta)
        dlog = torch.zeros_like(beta)  # d2 = 0: log(1 - d2) = 0
    mean, var = out if with_uncertainty else (out, None)
    contribution = values[3] * path
    torch.testing.assert_close(mean, contribution.sum().expand_as(mean))
    factor = torch.ones_like(beta)
    if with_uncertainty:
        torch.testing.assert_close(var, (q * contribution.square()).sum().expand_as(var))
        factor = factor + 0.8 * q * contribution
    bwd = r2splat_amd.rasterize_bwd if operation == "rasterize" else r2splat_amd.voxelize_bwd
    grads = bwd(ctx, torch.ones_like(mean), torch.full_like(var, 0.4) if with_uncertainty else None)
    torch.testing.assert_close(grads.densities, cameras * factor * path)
    torch.testing.assert_close(grads.beta, cameras * factor * contribution * dlog)
    if with_uncertainty:
        torch.testing.assert_close(grads.uncertainty, cameras * 0.4 * contribution.square())
    else:
        assert grads.uncertainty is None


@pytest.mark.parametrize("dtype", DTYPES)
def test_off_centre_beta_ellipsoid_voxels_by_hand(device, dtype):
    mean = (0.13, -0.21, 1.92)
    s = (0.7, 0.45, 0.3)
    rho, beta, q = 0.9, 2.5, 0.07
    shape = (3, 5, 4)  # (D, H, W)
    extent, center = (1.6, 1.1, 0.75), (0.05, -0.1, 2.0)
    t = lambda x: torch.tensor(x, device=device, dtype=dtype)
    (vol, var), _ = r2splat_amd.voxelize_fwd(t([mean]), t([[1.0, 0, 0, 0]]), t([s]), t([rho]), shape,
                                             volume_extent=extent, volume_center=center, beta=t([beta]),
                                             uncertainty=t([q]), eps=0.0)
    D, H, W = shape
    expect = torch.zeros(shape, dtype=torch.float64)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                p = [center[k] + ((i + 0.5) / n - 0.5) * extent[k] for k, (i, n) in enumerate(((x, W), (y, H), (z, D)))]
                d2 = sum((p[k] - mean[k]) ** 2 / s[k] ** 2 for k in range(3))
                a = rho * (1 - d2) ** beta if 0 <= d2 < 1 else 0.0
                expect[z, y, x] = a if a >= 1e-5 else 0.0
    assert 0 < int((expect > 0).sum()) < expect.numel()
    tol = 1e-6 if dtype == torch.float32 else 1e-13
    torch.testing.assert_close(vol.cpu().double(), expect, atol=tol, rtol=0)
    torch.testing.assert_close(var.cpu().double(), q * expect.square(), atol=tol, rtol=0)


@pytest.mark.parametrize("dtype", DTYPES)
def test_voxel_cutoffs_either_side_of_the_threshold(device, dtype):
    s = 0.5
    margin = 1e-4 if dtype == torch.float32 else 1e-9

    def one(d2, beta=None):
        t = lambda x: torch.tensor(x, device=device, dtype=dtype)
        out, _ = r2splat_amd.voxelize_fwd(t([[s * math.sqrt(d2), 0.0, 2.0]]), t([[1.0, 0, 0, 0]]), t([[s, s, s]]),
                                          t([1.0]), (1, 1, 1), volume_extent=(1, 1, 1), volume_center=(0, 0, 2),
                                          beta=None if beta is None else t([beta]), eps=0.0)
        return out.item()

    g_edge = -2 * math.log(1e-6)
    assert one(g_edge * (1 - margin)) > 0 and one(g_edge * (1 + margin)) == 0.0
    b_edge = 1 - 1e-5 ** 0.5  # beta 2: (1 - d2)^2 = 1e-5
    assert one(b_edge * (1 - margin), 2.0) > 0 and one(b_edge * (1 + margin), 2.0) == 0.0


@pytest.mark.parametrize("dtype", DTYPES)
def test_projection_radii_closed_form_and_culls(device, dtype):
    """Axis-aligned primitives, camera at the origin looking down +z, fx = 8,
    fy = 6, cx = 10, cy = 5 on a 20 x 10 image, z = 2. On the optical axis
    cov2d[0,0] = (fx sx / z)^2 and rx = ceil(3 sqrt(that + eps)); points far
    past each image edge and at or behind the near plane (z <= 0.01) get zero radii."""
    t = lambda x: torch.tensor(x, device=device, dtype=dtype)
    eps = 0.3
    scales = [[0.5, 0.25, 0.1], [0.3, 0.7, 0.1], [0.05, 0.05, 0.05]]
    means = [[0.0, 0.0, 2.0]] * 3 + [[-20.0, 0.0, 2.0], [20.0, 0.0, 2.0], [0.0, -20.0, 2.0], [0.0, 20.0, 2.0],
                                     [0.0, 0.0, 0.01], [0.0, 0.0, -1.0]]
    scales = scales + [[0.05, 0.05, 0.05]] * 6
    covars, _ = ops.parameter_transform_fwd(t([[1.0, 0, 0, 0]] * len(means)), t(scales))
    view = torch.eye(4, device=device, dtype=dtype)[None]
    K = t([[[8.0, 0, 10.0], [0, 6.0, 5.0], [0, 0, 1]]])
    _, _, radii = ops.world_to_image_projection_fwd(t(means), covars, view, K, 20, 10, eps, 3.0)
    radii = radii[0].cpu().tolist()
    for n in range(3):
        sx, sy = scales[n][0], scales[n][1]
        expect = [math.ceil(3 * math.sqrt((8 * sx / 2) ** 2 + eps)), math.ceil(3 * math.sqrt((6 * sy / 2) ** 2 + eps))]
        assert radii[n] == expect, (n, radii[n], expect)
    assert radii[3:] == [[0, 0]] * 6

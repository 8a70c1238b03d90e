// Per-primitive geometry: physical parameters -> covariance, world -> image
// projection (local perspective CT line integral), 2x2 conic + path factor, and
// world -> voxel mapping. One lane per primitive (or per camera x primitive).
//
// Semantics follow the reference operators (files cited per kernel); the 2x2
// and 3x3 covariance inversions run in double precision for every dtype.
#include "common.hpp"
#include "geometry_common.hpp"

namespace r2s {
template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
parameter_transform_fwd_kernel(const S* __restrict__ quats, const S* __restrict__ scales,
                               S* __restrict__ covars, S* __restrict__ max_scales, uint32_t N) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    S q[4], mag, norm, R[9], cov[6];
    unit_quat(quats + 4 * size_t(n), q, mag, norm);
    quat_to_rot(q, R);
    const S s[3] = {scales[3 * size_t(n)], scales[3 * size_t(n) + 1], scales[3 * size_t(n) + 2]};
    covariance_from(R, s, cov);
#pragma unroll
    for (int i = 0; i < 6; ++i) covars[6 * size_t(n) + i] = cov[i];
    max_scales[n] = max(s[0], max(s[1], s[2]));
}

template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
parameter_transform_bwd_kernel(const S* __restrict__ quats, const S* __restrict__ scales,
                               const S* __restrict__ v_covars, S* __restrict__ v_quats,
                               S* __restrict__ v_scales, uint32_t N) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    S q[4], mag, norm, R[9];
    unit_quat(quats + 4 * size_t(n), q, mag, norm);
    quat_to_rot(q, R);
    const S s[3] = {scales[3 * size_t(n)], scales[3 * size_t(n) + 1], scales[3 * size_t(n) + 2]};
    S vc[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) vc[i] = v_covars[6 * size_t(n) + i];
    S gq[4], vs[3];
    covariance_vjp(q, R, s, vc, gq, vs);
    normalize_vjp(q, mag, norm, gq, v_quats + 4 * size_t(n));
#pragma unroll
    for (int i = 0; i < 3; ++i) v_scales[3 * size_t(n) + i] = vs[i];
}

template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
projection_fwd_kernel(const S* __restrict__ means, const S* __restrict__ covars,
                      const S* __restrict__ viewmats, const S* __restrict__ Ks, uint32_t width,
                      uint32_t height, S eps, S clip, S* __restrict__ means2d,
                      S* __restrict__ cov2d, int32_t* __restrict__ radii, uint32_t C, uint32_t N) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = blockIdx.y;
    if (n >= N) return;
    const size_t idx = size_t(c) * N + n;
    Camera<S> cam;
    cam.load(viewmats + 16 * size_t(c), Ks + 9 * size_t(c));
    const S p[3] = {means[3 * size_t(n)], means[3 * size_t(n) + 1], means[3 * size_t(n) + 2]};
    S cov[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) cov[i] = covars[6 * size_t(n) + i];
    ProjState<S> st;
    const S W = S(width), H = S(height);
    if (!project_state(cam, p, cov, W, H, st)) {
        means2d[2 * idx] = S(0);
        means2d[2 * idx + 1] = S(0);
#pragma unroll
        for (int i = 0; i < 6; ++i) cov2d[6 * idx + i] = S(0);
        radii[2 * idx] = 0;
        radii[2 * idx + 1] = 0;
        return;
    }
    const S mx = cam.fx * st.x / st.z + cam.cx;
    const S my = cam.fy * st.y / st.z + cam.cy;
    means2d[2 * idx] = mx;
    means2d[2 * idx + 1] = my;
    S c2[6];
    projected_cov(st, c2);
#pragma unroll
    for (int i = 0; i < 6; ++i) cov2d[6 * idx + i] = c2[i];
    int32_t rx, ry;
    footprint_radii(c2, mx, my, eps, clip, W, H, rx, ry);
    radii[2 * idx] = rx;
    radii[2 * idx + 1] = ry;
}

// One lane per primitive, looping over cameras: the camera sum is done in a
// fixed order in registers, so the result is deterministic and needs no atomics.
template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
projection_bwd_kernel(const S* __restrict__ means, const S* __restrict__ covars,
                      const S* __restrict__ viewmats, const S* __restrict__ Ks, uint32_t width,
                      uint32_t height, const S* __restrict__ v_means2d,
                      const S* __restrict__ v_cov2d, S* __restrict__ v_means,
                      S* __restrict__ v_covars, uint32_t C, uint32_t N) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const S p[3] = {means[3 * size_t(n)], means[3 * size_t(n) + 1], means[3 * size_t(n) + 2]};
    S cov[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) cov[i] = covars[6 * size_t(n) + i];
    S vp[3] = {S(0), S(0), S(0)}, vcov[6] = {S(0), S(0), S(0), S(0), S(0), S(0)};
    const S W = S(width), H = S(height);
    for (uint32_t c = 0; c < C; ++c) {
        Camera<S> cam;
        cam.load(viewmats + 16 * size_t(c), Ks + 9 * size_t(c));
        ProjState<S> st;
        if (!project_state(cam, p, cov, W, H, st)) continue;
        const size_t idx = size_t(c) * N + n;
        const S vm[2] = {v_means2d[2 * idx], v_means2d[2 * idx + 1]};
        S vc[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) vc[i] = v_cov2d[6 * idx + i];
        projection_vjp(cam, st, vm, vc, vp, vcov);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) v_means[3 * size_t(n) + i] = vp[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) v_covars[6 * size_t(n) + i] = vcov[i];
}

template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
conics_mu_fwd_kernel(const S* __restrict__ cov2d, double eps, S* __restrict__ conics,
                     S* __restrict__ mus, uint32_t total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    double c[6], P[3], mu;
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = double(cov2d[6 * size_t(i) + k]);
    // eps is rounded to the working precision first, as the reference does.
    conic_mu_f64(c, double(S(eps)), P, mu);
    conics[3 * size_t(i)] = S(P[0]);
    conics[3 * size_t(i) + 1] = S(P[1]);
    conics[3 * size_t(i) + 2] = S(P[2]);
    mus[i] = S(mu);
}

template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
conics_mu_bwd_kernel(const S* __restrict__ cov2d, const S* __restrict__ conics,
                     const S* __restrict__ mus, const S* __restrict__ v_conics,
                     const S* __restrict__ v_mus, S* __restrict__ v_cov2d, uint32_t total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    double c[6], P[3], vP[3], g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = double(cov2d[6 * size_t(i) + k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        P[k] = double(conics[3 * size_t(i) + k]);
        vP[k] = double(v_conics[3 * size_t(i) + k]);
    }
    conic_mu_vjp_f64(c, P, double(mus[i]), vP, double(v_mus[i]), g);
#pragma unroll
    for (int k = 0; k < 6; ++k) v_cov2d[6 * size_t(i) + k] = S(g[k]);
}

template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
volume_mapping_fwd_kernel(const S* __restrict__ means, const S* __restrict__ covars,
                          const S* __restrict__ ranges, const S* __restrict__ centers,
                          uint32_t dx, uint32_t dy, uint32_t dz, S eps, S clip,
                          S* __restrict__ mapped, S* __restrict__ conics,
                          int32_t* __restrict__ radii, uint32_t N) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const S D[3] = {S(dx), S(dy), S(dz)};
    const S r[3] = {D[0] / ranges[0], D[1] / ranges[1], D[2] / ranges[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        mapped[3 * size_t(n) + k] = (means[3 * size_t(n) + k] - centers[k]) * r[k] + D[k] * S(0.5);
    const S* s = covars + 6 * size_t(n);
    // Voxel-frame covariance in the working precision (its diagonal sets the radii).
    const S sv[6] = {r[0] * r[0] * s[0] + eps, r[0] * r[1] * s[1], r[0] * r[2] * s[2],
                     r[1] * r[1] * s[3] + eps, r[1] * r[2] * s[4], r[2] * r[2] * s[5] + eps};
    double a[6], inv[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = double(sv[k]);
    const bool ok = spd3_inverse(a, inv);
#pragma unroll
    for (int k = 0; k < 6; ++k) conics[6 * size_t(n) + k] = S(inv[k]);
    int32_t rr[3] = {0, 0, 0};
    if (ok) {
        const S rad[3] = {ceil(clip * sqrt(max(S(0), sv[0]))), ceil(clip * sqrt(max(S(0), sv[3]))),
                          ceil(clip * sqrt(max(S(0), sv[5])))};
        bool good = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) good = good && isfinite(rad[k]) && rad[k] > S(0);
        if (good) {
#pragma unroll
            for (int k = 0; k < 3; ++k) rr[k] = static_cast<int32_t>(rad[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) radii[3 * size_t(n) + k] = rr[k];
}

// Inverse VJP: v_Sv = -P V P with V the upper-triangular packed adjoint.
template <typename S>
__global__ void __launch_bounds__(kGeomThreads)
volume_mapping_bwd_kernel(const S* __restrict__ covars, const S* __restrict__ ranges,
                          uint32_t dx, uint32_t dy, uint32_t dz, S eps,
                          const S* __restrict__ v_mapped, const S* __restrict__ v_conics,
                          S* __restrict__ v_means, S* __restrict__ v_covars, uint32_t N) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const S r[3] = {S(dx) / ranges[0], S(dy) / ranges[1], S(dz) / ranges[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) v_means[3 * size_t(n) + k] = v_mapped[3 * size_t(n) + k] * r[k];
    const S* s = covars + 6 * size_t(n);
    const S sv[6] = {r[0] * r[0] * s[0] + eps, r[0] * r[1] * s[1], r[0] * r[2] * s[2],
                     r[1] * r[1] * s[3] + eps, r[1] * r[2] * s[4], r[2] * r[2] * s[5] + eps};
    double a[6], P6[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = double(sv[k]);
    S* out = v_covars + 6 * size_t(n);
    if (!spd3_inverse(a, P6)) {
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k] = S(0);
        return;
    }
    const double P[9] = {P6[0], P6[1], P6[2], P6[1], P6[3], P6[4], P6[2], P6[4], P6[5]};
    const S* vc = v_conics + 6 * size_t(n);
    const double V[9] = {double(vc[0]), double(vc[1]), double(vc[2]), 0.0, double(vc[3]),
                         double(vc[4]), 0.0, 0.0, double(vc[5])};
    double VP[9], G[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            VP[3 * i + j] = V[3 * i] * P[j] + V[3 * i + 1] * P[3 + j] + V[3 * i + 2] * P[6 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            G[3 * i + j] = -(P[3 * i] * VP[j] + P[3 * i + 1] * VP[3 + j] + P[3 * i + 2] * VP[6 + j]);
    const double rd[3] = {double(r[0]), double(r[1]), double(r[2])};
    out[0] = S(G[0] * rd[0] * rd[0]);
    out[1] = S((G[1] + G[3]) * rd[0] * rd[1]);
    out[2] = S((G[2] + G[6]) * rd[0] * rd[2]);
    out[3] = S(G[4] * rd[1] * rd[1]);
    out[4] = S((G[5] + G[7]) * rd[1] * rd[2]);
    out[5] = S(G[8] * rd[2] * rd[2]);
}

}  // namespace r2s

using namespace r2s;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int r2s_parameter_transform_fwd(int dtype, const void* quats, const void* scales,
                                           void* covars, void* max_scales, uint32_t N,
                                           r2s_stream_t stream) {
    if (N == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        parameter_transform_fwd_kernel<S><<<ceil_div(N, kGeomThreads), kGeomThreads, 0, stream>>>(
            (const S*)quats, (const S*)scales, (S*)covars, (S*)max_scales, N);
        return check_launch("parameter_transform_fwd");
    });
}

extern "C" int r2s_parameter_transform_bwd(int dtype, const void* quats, const void* scales,
                                           const void* v_covars, void* v_quats, void* v_scales,
                                           uint32_t N, r2s_stream_t stream) {
    if (N == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        parameter_transform_bwd_kernel<S><<<ceil_div(N, kGeomThreads), kGeomThreads, 0, stream>>>(
            (const S*)quats, (const S*)scales, (const S*)v_covars, (S*)v_quats, (S*)v_scales, N);
        return check_launch("parameter_transform_bwd");
    });
}

extern "C" int r2s_world_to_image_projection_fwd(int dtype, const void* means, const void* covars,
                                                 const void* viewmats, const void* Ks,
                                                 uint32_t width, uint32_t height, float eps,
                                                 float radius_clip, void* means2d, void* cov2d,
                                                 int32_t* radii, uint32_t C, uint32_t N,
                                                 r2s_stream_t stream) {
    R2S_REQUIRE(radius_clip > 0.0f, "radius_clip must be positive, got %g", double(radius_clip));
    R2S_REQUIRE(C <= 65535u, "camera count %u exceeds the grid limit", C);
    if (C == 0 || N == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        dim3 grid(ceil_div(N, kGeomThreads), C);
        projection_fwd_kernel<S><<<grid, kGeomThreads, 0, stream>>>(
            (const S*)means, (const S*)covars, (const S*)viewmats, (const S*)Ks, width, height,
            S(eps), S(radius_clip), (S*)means2d, (S*)cov2d, radii, C, N);
        return check_launch("world_to_image_projection_fwd");
    });
}

extern "C" int r2s_world_to_image_projection_bwd(int dtype, const void* means, const void* covars,
                                                 const void* viewmats, const void* Ks,
                                                 uint32_t width, uint32_t height,
                                                 const void* v_means2d, const void* v_cov2d,
                                                 void* v_means, void* v_covars, uint32_t C,
                                                 uint32_t N, r2s_stream_t stream) {
    if (N == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        projection_bwd_kernel<S><<<ceil_div(N, kGeomThreads), kGeomThreads, 0, stream>>>(
            (const S*)means, (const S*)covars, (const S*)viewmats, (const S*)Ks, width, height,
            (const S*)v_means2d, (const S*)v_cov2d, (S*)v_means, (S*)v_covars, C, N);
        return check_launch("world_to_image_projection_bwd");
    });
}

extern "C" int r2s_conics_mu_fwd(int dtype, const void* cov2d, double eps, void* conics, void* mus,
                                 uint32_t total, r2s_stream_t stream) {
    if (total == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        conics_mu_fwd_kernel<S><<<ceil_div(total, kGeomThreads), kGeomThreads, 0, stream>>>(
            (const S*)cov2d, eps, (S*)conics, (S*)mus, total);
        return check_launch("conics_mu_fwd");
    });
}

extern "C" int r2s_conics_mu_bwd(int dtype, const void* cov2d, double eps, const void* conics,
                                 const void* mus, const void* v_conics, const void* v_mus,
                                 void* v_cov2d, uint32_t total, r2s_stream_t stream) {
    (void)eps;  // the VJP only needs the saved conics and mu
    if (total == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        conics_mu_bwd_kernel<S><<<ceil_div(total, kGeomThreads), kGeomThreads, 0, stream>>>(
            (const S*)cov2d, (const S*)conics, (const S*)mus, (const S*)v_conics, (const S*)v_mus,
            (S*)v_cov2d, total);
        return check_launch("conics_mu_bwd");
    });
}

extern "C" int r2s_world_to_volume_mapping_fwd(int dtype, const void* means, const void* covars,
                                               const void* max_scales, const void* ranges,
                                               const void* centers, uint32_t dim_x,
                                               uint32_t dim_y, uint32_t dim_z, double eps3d,
                                               double radius_clip, void* means3d_v, void* conics,
                                               int32_t* radii, uint32_t N, r2s_stream_t stream) {
    (void)max_scales;
    if (N == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        volume_mapping_fwd_kernel<S><<<ceil_div(N, kGeomThreads), kGeomThreads, 0, stream>>>(
            (const S*)means, (const S*)covars, (const S*)ranges, (const S*)centers, dim_x, dim_y,
            dim_z, S(eps3d), S(radius_clip), (S*)means3d_v, (S*)conics, radii, N);
        return check_launch("world_to_volume_mapping_fwd");
    });
}

extern "C" int r2s_world_to_volume_mapping_bwd(int dtype, const void* means, const void* covars,
                                               const void* ranges, const void* centers,
                                               uint32_t dim_x, uint32_t dim_y, uint32_t dim_z,
                                               double eps3d, const void* v_means3d_v,
                                               const void* v_conics, void* v_means,
                                               void* v_covars, uint32_t N, r2s_stream_t stream) {
    (void)means;
    (void)centers;
    if (N == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        volume_mapping_bwd_kernel<S><<<ceil_div(N, kGeomThreads), kGeomThreads, 0, stream>>>(
            (const S*)covars, (const S*)ranges, dim_x, dim_y, dim_z, S(eps3d),
            (const S*)v_means3d_v, (const S*)v_conics, (S*)v_means, (S*)v_covars, N);
        return check_launch("world_to_volume_mapping_bwd");
    });
}

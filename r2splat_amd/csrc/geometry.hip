// Per-primitive geometry: physical parameters -> covariance, world -> image
// projection (local perspective CT line integral), 2x2 conic + path factor, and
// world -> voxel mapping. One lane per primitive (or per camera x primitive).
//
// Semantics follow the reference operators (files cited per kernel); the 2x2
// and 3x3 covariance inversions run in double precision for every dtype.
#include "common.hpp"

namespace r2s {

constexpr int kGeomThreads = 256;

// ---------------------------------------------------------------------------
// Quaternion helpers (r2splat/cuda/include/ParameterTransform.cuh:44-74).
// Normalisation first divides by max|q_i| so that very large or subnormal
// raw quaternions neither overflow nor flush: q_hat = (q / m) / |q / m|.
// Division and sqrt are IEEE correctly rounded (no fast-math in this file).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float max_abs4(const float* q) {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) m = max(m, __float_as_uint(q[i]) & 0x7fffffffu);
    return __uint_as_float(m);
}
__device__ __forceinline__ double max_abs4(const double* q) {
    return fmax(fmax(fabs(q[0]), fabs(q[1])), fmax(fabs(q[2]), fabs(q[3])));
}

template <typename S>
__device__ __forceinline__ void unit_quat(const S* raw, S q[4], S& mag, S& norm) {
    mag = max_abs4(raw);
    S ss = S(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        q[i] = raw[i] / mag;
        ss += q[i] * q[i];
    }
    norm = sqrt(ss);
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = q[i] / norm;
}

// Rotation matrix of a unit quaternion (w, x, y, z), row-major.
template <typename S>
__device__ __forceinline__ void quat_to_rot(const S q[4], S R[9]) {
    const S w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = S(1) - S(2) * (y * y + z * z);
    R[1] = S(2) * (x * y - w * z);
    R[2] = S(2) * (x * z + w * y);
    R[3] = S(2) * (x * y + w * z);
    R[4] = S(1) - S(2) * (x * x + z * z);
    R[5] = S(2) * (y * z - w * x);
    R[6] = S(2) * (x * z - w * y);
    R[7] = S(2) * (y * z + w * x);
    R[8] = S(1) - S(2) * (x * x + y * y);
}

// covars = R diag(s^2) R^T, packed (00,01,02,11,12,22).
template <typename S>
__device__ __forceinline__ void covariance_from(const S R[9], const S s[3], S cov[6]) {
    const S s2[3] = {s[0] * s[0], s[1] * s[1], s[2] * s[2]};
    int k = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            cov[k++] = R[3 * i + 0] * R[3 * j + 0] * s2[0] + R[3 * i + 1] * R[3 * j + 1] * s2[1] +
                       R[3 * i + 2] * R[3 * j + 2] * s2[2];
        }
}

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

// VJP of covars(q, s) (r2splat/cuda/csrc/ParameterTransformBwd.cu:10-77).
// With M = R diag(s): Sigma = M M^T, so dL/dM = (G + G^T) M = 2 G M for the
// symmetric G whose off-diagonals carry half of each packed off-diagonal.
template <typename S>
__device__ __forceinline__ void covariance_vjp(const S q[4], const S R[9], const S s[3],
                                               const S vc[6], S v_q_unit[4], S v_s[3]) {
    const S G[9] = {vc[0],        S(0.5) * vc[1], S(0.5) * vc[2],
                    S(0.5) * vc[1], vc[3],        S(0.5) * vc[4],
                    S(0.5) * vc[2], S(0.5) * vc[4], vc[5]};
    S vR[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        S acc = S(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const S gm = S(2) * s[j] *
                         (G[3 * i + 0] * R[0 + j] + G[3 * i + 1] * R[3 + j] + G[3 * i + 2] * R[6 + j]);
            vR[3 * i + j] = gm * s[j];
            acc += R[3 * i + j] * gm;
        }
        v_s[j] = acc;
    }
    const S w = q[0], x = q[1], y = q[2], z = q[3];
    v_q_unit[0] = S(2) * (x * (vR[7] - vR[5]) + y * (vR[2] - vR[6]) + z * (vR[3] - vR[1]));
    v_q_unit[1] = S(2) * (-S(2) * x * (vR[4] + vR[8]) + y * (vR[1] + vR[3]) + z * (vR[2] + vR[6]) +
                          w * (vR[7] - vR[5]));
    v_q_unit[2] = S(2) * (x * (vR[1] + vR[3]) - S(2) * y * (vR[0] + vR[8]) + z * (vR[5] + vR[7]) +
                          w * (vR[2] - vR[6]));
    v_q_unit[3] = S(2) * (x * (vR[2] + vR[6]) + y * (vR[5] + vR[7]) - S(2) * z * (vR[0] + vR[4]) +
                          w * (vR[3] - vR[1]));
}

// Normalisation VJP: d(q/|q|)^T g = (g - q_hat (q_hat . g)) / |q|, with |q|
// applied as two sequential divisions (by the scaled norm, then by max|q_i|).
template <typename S>
__device__ __forceinline__ void normalize_vjp(const S q[4], S mag, S norm, const S g[4],
                                              S* __restrict__ out) {
    const S qg = q[0] * g[0] + q[1] * g[1] + q[2] * g[2] + q[3] * g[3];
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = ((g[i] - q[i] * qg) / norm) / mag;
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

// ---------------------------------------------------------------------------
// World -> image projection (WorldToImageProjectionFwd.cu:73-283).
// ---------------------------------------------------------------------------
template <typename S>
struct Camera {
    S R[9], t[3], fx, fy, cx, cy;
    __device__ __forceinline__ void load(const S* __restrict__ viewmat, const S* __restrict__ K) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) R[3 * i + j] = viewmat[4 * i + j];
            t[i] = viewmat[4 * i + 3];
        }
        fx = K[0];
        fy = K[4];
        cx = K[2];
        cy = K[5];
    }
};

// Projection intermediates shared by the forward and the VJP.
template <typename S>
struct ProjState {
    S x, y, z;          // camera-frame mean
    S tx, ty, l;        // FOV-clamped ray point and its length
    bool clip_x, clip_y;
    S J[9];             // rows: d(pixel x), d(pixel y), unit ray direction
    S Sc[6];            // camera-frame covariance, packed
};

// Returns false when the primitive lies at depth z <= 0.01 (culled).
template <typename S>
__device__ __forceinline__ bool project_state(const Camera<S>& cam, const S p[3], const S cov[6],
                                              S width, S height, ProjState<S>& st) {
    const S* R = cam.R;
    st.x = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + cam.t[0];
    st.y = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + cam.t[1];
    st.z = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + cam.t[2];
    if (!(st.z > S(0.01))) return false;
    const S rz = S(1) / st.z;
    // Field-of-view clamp with a 15% image margin on each side.
    const S lim_xp = (S(1.15) * width - cam.cx) / cam.fx;
    const S lim_xn = (cam.cx + S(0.15) * width) / cam.fx;
    const S lim_yp = (S(1.15) * height - cam.cy) / cam.fy;
    const S lim_yn = (cam.cy + S(0.15) * height) / cam.fy;
    const S u = st.x * rz, v = st.y * rz;
    st.clip_x = (u > lim_xp) || (u < -lim_xn);
    st.clip_y = (v > lim_yp) || (v < -lim_yn);
    st.tx = st.z * min(lim_xp, max(-lim_xn, u));
    st.ty = st.z * min(lim_yp, max(-lim_yn, v));
    st.l = sqrt(st.tx * st.tx + st.ty * st.ty + st.z * st.z);
    const S rl = S(1) / st.l;
    st.J[0] = cam.fx * rz;
    st.J[1] = S(0);
    st.J[2] = -cam.fx * st.tx * rz * rz;
    st.J[3] = S(0);
    st.J[4] = cam.fy * rz;
    st.J[5] = -cam.fy * st.ty * rz * rz;
    st.J[6] = st.tx * rl;
    st.J[7] = st.ty * rl;
    st.J[8] = st.z * rl;
    // Sc = R Sigma R^T
    const S Sg[9] = {cov[0], cov[1], cov[2], cov[1], cov[3], cov[4], cov[2], cov[4], cov[5]};
    S RS[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            RS[3 * i + j] = R[3 * i + 0] * Sg[0 + j] + R[3 * i + 1] * Sg[3 + j] + R[3 * i + 2] * Sg[6 + j];
    int k = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j)
            st.Sc[k++] = RS[3 * i + 0] * R[3 * j + 0] + RS[3 * i + 1] * R[3 * j + 1] + RS[3 * i + 2] * R[3 * j + 2];
    return true;
}

// cov2d = J Sc J^T (packed upper triangle).
template <typename S>
__device__ __forceinline__ void projected_cov(const ProjState<S>& st, S out[6]) {
    const S Sc[9] = {st.Sc[0], st.Sc[1], st.Sc[2], st.Sc[1], st.Sc[3], st.Sc[4], st.Sc[2], st.Sc[4], st.Sc[5]};
    S JS[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            JS[3 * i + j] = st.J[3 * i + 0] * Sc[0 + j] + st.J[3 * i + 1] * Sc[3 + j] + st.J[3 * i + 2] * Sc[6 + j];
    int k = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j)
            out[k++] = JS[3 * i + 0] * st.J[3 * j + 0] + JS[3 * i + 1] * st.J[3 * j + 1] + JS[3 * i + 2] * st.J[3 * j + 2];
}

// Axis-aligned pixel radii of the blurred 2x2 footprint (non-differentiable),
// computed in the working precision exactly as the reference does
// (WorldToImageProjectionFwd.cu:250-282).
template <typename S>
__device__ __forceinline__ void footprint_radii(const S c2[6], S mx, S my, S eps, S clip,
                                                S width, S height, int32_t& rx, int32_t& ry) {
    rx = ry = 0;
    const S a = c2[0] + eps, d = c2[3] + eps, b = c2[1];
    if (!(a * d - b * b > S(0))) return;
    const S fx = ceil(clip * sqrt(max(S(0), a)));
    const S fy = ceil(clip * sqrt(max(S(0), d)));
    if (!isfinite(fx) || !isfinite(fy) || !(fx > S(0)) || !(fy > S(0))) return;
    if (mx + fx <= S(0) || mx - fx >= width || my + fy <= S(0) || my - fy >= height) return;
    rx = static_cast<int32_t>(fx);
    ry = static_cast<int32_t>(fy);
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

// VJP of (means2d, cov2d) w.r.t. the camera-frame mean and the world covariance
// for one (camera, primitive) pair (WorldToImageProjectionBwd.cu:67-554).
// vm: d/d means2d; vc: d/d packed cov2d. Accumulates into v_p (world) and v_cov.
template <typename S>
__device__ __forceinline__ void projection_vjp(const Camera<S>& cam, const ProjState<S>& st,
                                               const S vm[2], const S vc[6], S v_p[3],
                                               S v_cov[6]) {
    const S z = st.z, rz = S(1) / z, rz2 = rz * rz, rz3 = rz2 * rz;
    const S rl = S(1) / st.l, rl3 = rl * rl * rl;
    const S tx = st.tx, ty = st.ty;
    // G: upper-triangular adjoint of the packed symmetric output.
    const S G[9] = {vc[0], vc[1], vc[2], S(0), vc[3], vc[4], S(0), S(0), vc[5]};
    const S* J = st.J;
    // v_Sc = J^T G J (not symmetric).
    S GJ[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            GJ[3 * i + j] = G[3 * i + 0] * J[0 + j] + G[3 * i + 1] * J[3 + j] + G[3 * i + 2] * J[6 + j];
    S vSc[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vSc[3 * i + j] = J[0 + i] * GJ[0 + j] + J[3 + i] * GJ[3 + j] + J[6 + i] * GJ[6 + j];
    // v_Sigma = R^T v_Sc R; packed: diagonal once, off-diagonals as a_ij + a_ji.
    const S* R = cam.R;
    S T1[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            T1[3 * i + j] = vSc[3 * i + 0] * R[0 + j] + vSc[3 * i + 1] * R[3 + j] + vSc[3 * i + 2] * R[6 + j];
    S vS[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vS[3 * i + j] = R[0 + i] * T1[0 + j] + R[3 + i] * T1[3 + j] + R[6 + i] * T1[6 + j];
    v_cov[0] += vS[0];
    v_cov[1] += vS[1] + vS[3];
    v_cov[2] += vS[2] + vS[6];
    v_cov[3] += vS[4];
    v_cov[4] += vS[5] + vS[7];
    v_cov[5] += vS[8];
    // v_J = (G + G^T) J Sc.
    const S Sc[9] = {st.Sc[0], st.Sc[1], st.Sc[2], st.Sc[1], st.Sc[3], st.Sc[4], st.Sc[2], st.Sc[4], st.Sc[5]};
    S JSc[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            JSc[3 * i + j] = J[3 * i + 0] * Sc[0 + j] + J[3 * i + 1] * Sc[3 + j] + J[3 * i + 2] * Sc[6 + j];
    const S GG[9] = {S(2) * vc[0], vc[1], vc[2], vc[1], S(2) * vc[3], vc[4], vc[2], vc[4], S(2) * vc[5]};
    S vJ[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vJ[3 * i + j] = GG[3 * i + 0] * JSc[0 + j] + GG[3 * i + 1] * JSc[3 + j] + GG[3 * i + 2] * JSc[6 + j];
    // Partial derivatives of J w.r.t. (tx, ty, z).
    const S v_tx = vJ[2] * (-cam.fx * rz2) + vJ[6] * (rl - tx * tx * rl3) + vJ[7] * (-tx * ty * rl3) +
                   vJ[8] * (-z * tx * rl3);
    const S v_ty = vJ[5] * (-cam.fy * rz2) + vJ[6] * (-tx * ty * rl3) + vJ[7] * (rl - ty * ty * rl3) +
                   vJ[8] * (-z * ty * rl3);
    S v_z = vJ[0] * (-cam.fx * rz2) + vJ[2] * (S(2) * cam.fx * tx * rz3) + vJ[4] * (-cam.fy * rz2) +
            vJ[5] * (S(2) * cam.fy * ty * rz3) + vJ[6] * (-tx * z * rl3) + vJ[7] * (-ty * z * rl3) +
            vJ[8] * (rl - z * z * rl3);
    S v_x = S(0), v_y = S(0);
    // tx = x when unclamped, z * limit when clamped (then dtx/dz = tx / z).
    if (st.clip_x) v_z += v_tx * tx * rz; else v_x += v_tx;
    if (st.clip_y) v_z += v_ty * ty * rz; else v_y += v_ty;
    // means2d = (fx x / z + cx, fy y / z + cy).
    v_x += cam.fx * rz * vm[0];
    v_y += cam.fy * rz * vm[1];
    v_z -= (cam.fx * st.x * vm[0] + cam.fy * st.y * vm[1]) * rz2;
    // world mean: v_p = R^T v_pc.
    v_p[0] += R[0] * v_x + R[3] * v_y + R[6] * v_z;
    v_p[1] += R[1] * v_x + R[4] * v_y + R[7] * v_z;
    v_p[2] += R[2] * v_x + R[5] * v_y + R[8] * v_z;
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

// ---------------------------------------------------------------------------
// Conic + path factor (ConicsMuFwd.cu:35-104, ConicsMuBwd.cu:44-129), double
// precision internally. Degenerate 2x2 blocks give zero conics and mu.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool conic_mu_f64(const double c[6], double eps, double P[3], double& mu) {
    const double a = c[0] + eps, b = c[1], d = c[3] + eps;
    const double scale = fmax(fmax(fabs(a), fabs(d)), fabs(b));
    const double an = a / scale, bn = b / scale, dn = d / scale;
    const double det = an * dn - bn * bn;
    P[0] = (dn / det) / scale;
    P[1] = (-bn / det) / scale;
    P[2] = (an / det) / scale;
    if (!(a > 0.0 && d > 0.0 && det > 0.0) || !isfinite(P[0]) || !isfinite(P[1]) || !isfinite(P[2])) {
        P[0] = P[1] = P[2] = 0.0;
        mu = 0.0;
        return false;
    }
    const double w0 = P[0] * c[2] + P[1] * c[4];
    const double w1 = P[1] * c[2] + P[2] * c[4];
    const double s = c[5] + eps - c[2] * w0 - c[4] * w1;
    mu = (s > 0.0 && isfinite(s)) ? ::sqrt(2.0 * M_PI * s) : 0.0;
    return true;
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

// d(conics, mu)/d(cov2d): dP = -P dA P and ds = w dA w^T - 2 w dc + df with
// w = P [c, e]^T, mu = sqrt(2 pi s), dmu/ds = pi / mu.
__device__ __forceinline__ void conic_mu_vjp_f64(const double c[6], const double P[3], double mu,
                                                 const double vP[3], double v_mu, double g[6]) {
    const double u11 = P[0], u12 = P[1], u22 = P[2];
    for (int k = 0; k < 6; ++k) g[k] = 0.0;
    if (u11 == 0.0 && u12 == 0.0 && u22 == 0.0) return;
    g[0] = -(vP[0] * u11 * u11 + vP[1] * u11 * u12 + vP[2] * u12 * u12);
    g[1] = -(2.0 * vP[0] * u11 * u12 + vP[1] * (u11 * u22 + u12 * u12) + 2.0 * vP[2] * u12 * u22);
    g[3] = -(vP[0] * u12 * u12 + vP[1] * u12 * u22 + vP[2] * u22 * u22);
    if (mu > 0.0) {
        const double w0 = u11 * c[2] + u12 * c[4];
        const double w1 = u12 * c[2] + u22 * c[4];
        const double vs = (M_PI / mu) * v_mu;
        g[0] += vs * w0 * w0;
        g[1] += 2.0 * vs * w0 * w1;
        g[2] -= 2.0 * vs * w0;
        g[3] += vs * w1 * w1;
        g[4] -= 2.0 * vs * w1;
        g[5] += vs;
    }
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

// ---------------------------------------------------------------------------
// World -> volume mapping (WorldToVolumeMappingFwd.cu:57-157,
// WorldToVolumeMappingBwd.cu:51-167, CovarianceMath.cuh:10-68).
// ---------------------------------------------------------------------------

// Inverse of a symmetric positive-definite 3x3 (packed), by Cholesky of the
// matrix scaled by its largest diagonal entry, so the validity test does not
// depend on the absolute scale. Returns false (and zeros) for singular,
// indefinite or non-finite input.
__device__ __forceinline__ bool spd3_inverse(const double a_in[6], double inv[6]) {
    for (int k = 0; k < 6; ++k) inv[k] = 0.0;
    for (int k = 0; k < 6; ++k)
        if (!isfinite(a_in[k])) return false;
    if (!(a_in[0] > 0.0) || !(a_in[3] > 0.0) || !(a_in[5] > 0.0)) return false;
    const double sc = fmax(a_in[0], fmax(a_in[3], a_in[5]));
    const double a00 = a_in[0] / sc, a01 = a_in[1] / sc, a02 = a_in[2] / sc;
    const double a11 = a_in[3] / sc, a12 = a_in[4] / sc, a22 = a_in[5] / sc;
    const double l00 = ::sqrt(a00);
    const double l10 = a01 / l00, l20 = a02 / l00;
    const double p11 = a11 - l10 * l10;
    if (!(p11 > 0.0)) return false;
    const double l11 = ::sqrt(p11);
    const double l21 = (a12 - l20 * l10) / l11;
    const double p22 = a22 - l20 * l20 - l21 * l21;
    if (!(p22 > 0.0)) return false;
    const double l22 = ::sqrt(p22);
    // K = L^-1 (lower triangular), A^-1 = K^T K / sc.
    const double k00 = 1.0 / l00, k11 = 1.0 / l11, k22 = 1.0 / l22;
    const double k10 = -l10 * k00 * k11;
    const double k21 = -l21 * k11 * k22;
    const double k20 = -(l20 * k00 + l21 * k10) * k22;
    const double r[6] = {(k00 * k00 + k10 * k10 + k20 * k20) / sc, (k10 * k11 + k20 * k21) / sc,
                         (k20 * k22) / sc, (k11 * k11 + k21 * k21) / sc, (k21 * k22) / sc,
                         (k22 * k22) / sc};
    for (int k = 0; k < 6; ++k)
        if (!isfinite(r[k])) return false;
    for (int k = 0; k < 6; ++k) inv[k] = r[k];
    return true;
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

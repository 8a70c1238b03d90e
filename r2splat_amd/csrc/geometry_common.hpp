// Device-side geometry shared by the operator kernels and the fused pipeline:
// quaternion -> covariance, local-perspective projection and its VJP, the
// double-precision 2x2 conic / path factor and the SPD 3x3 inverse.
#pragma once

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

// ---------------------------------------------------------------------------
// World -> image projection (WorldToImageProjectionFwd.cu:73-283).
// ---------------------------------------------------------------------------
template <typename S>
struct Camera {
    S R[9], t[3], fx, fy, cx, cy;
    // from a viewmat [4,4] and K [3,3] of any floating type (float32 cameras
    // are exact in double)
    template <typename T>
    __device__ __forceinline__ void load(const T* __restrict__ viewmat, const T* __restrict__ K) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) R[3 * i + j] = S(viewmat[4 * i + j]);
            t[i] = S(viewmat[4 * i + 3]);
        }
        fx = S(K[0]);
        fy = S(K[4]);
        cx = S(K[2]);
        cy = S(K[5]);
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

// Field-of-view clamp limits of a camera (15% image margin on each side), in
// normalised image coordinates: x in [-xn, xp], y in [-yn, yp].
template <typename S>
struct FovLimits {
    S xp, xn, yp, yn;
};

template <typename S>
__device__ __forceinline__ FovLimits<S> fov_limits(const Camera<S>& cam, S width, S height) {
    FovLimits<S> lim;
    lim.xp = (S(1.15) * width - cam.cx) / cam.fx;
    lim.xn = (cam.cx + S(0.15) * width) / cam.fx;
    lim.yp = (S(1.15) * height - cam.cy) / cam.fy;
    lim.yn = (cam.cy + S(0.15) * height) / cam.fy;
    return lim;
}

// Returns false when the primitive lies at depth z <= 0.01 (culled). `fov` is
// fov_limits(cam, width, height) (computed once per camera by callers that
// project many primitives per camera).
template <typename S>
__device__ __forceinline__ bool project_state(const Camera<S>& cam, const FovLimits<S>& fov, const S p[3],
                                              const S cov[6], ProjState<S>& st) {
    const S* R = cam.R;
    st.x = R[0] * p[0] + R[1] * p[1] + R[2] * p[2] + cam.t[0];
    st.y = R[3] * p[0] + R[4] * p[1] + R[5] * p[2] + cam.t[1];
    st.z = R[6] * p[0] + R[7] * p[1] + R[8] * p[2] + cam.t[2];
    if (!(st.z > S(0.01))) return false;
    const S rz = S(1) / st.z;
    const S lim_xp = fov.xp, lim_xn = fov.xn, lim_yp = fov.yp, lim_yn = fov.yn;
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

template <typename S>
__device__ __forceinline__ bool project_state(const Camera<S>& cam, const S p[3], const S cov[6],
                                              S width, S height, ProjState<S>& st) {
    return project_state(cam, fov_limits(cam, width, height), p, cov, st);
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

// conic_mu_f64 with two divisions instead of nine (the reciprocal of the
// scale and of det x scale, then products): the same values to within a few
// double ulps, for the fused kernels that evaluate it per (camera, primitive).
__device__ __forceinline__ bool conic_mu_fast_f64(const double c[6], double eps, double P[3], double& mu) {
    const double a = c[0] + eps, b = c[1], d = c[3] + eps;
    const double scale = fmax(fmax(fabs(a), fabs(d)), fabs(b));
    const double rs = 1.0 / scale;
    const double an = a * rs, bn = b * rs, dn = d * rs;
    const double det = an * dn - bn * bn;
    const double k = rs / det;
    P[0] = dn * k;
    P[1] = -bn * k;
    P[2] = an * k;
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

}  // namespace r2s

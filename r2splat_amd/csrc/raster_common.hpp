// Per-pixel profile evaluation and VJP for the 2D rasterizer, plus fill helpers.
#pragma once

#include "common.hpp"

namespace r2s {

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, size_t n, T value) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t k = i; k < n; k += stride) p[k] = value;
}

// Clears n elements with a kernel (graph-capture safe, unlike hipMemsetAsync).
template <typename T>
inline int fill_zero(T* p, size_t n, hipStream_t stream) {
    if (p == nullptr || n == 0) return R2S_OK;
    const uint32_t blocks = uint32_t(std::min<size_t>((n + 255) / 256, 8192));
    fill_kernel<T><<<blocks, 256, 0, stream>>>(p, n, T(0));
    return check_launch("fill_zero");
}

template <typename S>
inline int zero_grads_2d(size_t CN, void* v_abs, void* v_xy, void* v_conic, void* v_opac, void* v_mu,
                         void* v_beta, hipStream_t stream) {
    if (int rc = fill_zero<S>((S*)v_abs, 2 * CN, stream)) return rc;
    if (int rc = fill_zero<S>((S*)v_xy, 2 * CN, stream)) return rc;
    if (int rc = fill_zero<S>((S*)v_conic, 3 * CN, stream)) return rc;
    if (int rc = fill_zero<S>((S*)v_opac, CN, stream)) return rc;
    if (int rc = fill_zero<S>((S*)v_mu, CN, stream)) return rc;
    return fill_zero<S>((S*)v_beta, CN, stream);
}

template <typename S>
struct Prim2D {
    S mx, my, a, b, c, opac, mu, beta;
};

// Gaussian: delta = pixel - mean, sigma = 1/2 delta^T P delta, alpha = o mu e^-sigma,
//           skipped when sigma < 0 or |alpha| < 1e-5.
// Beta:     delta = mean - pixel, d2 = delta^T P delta, alpha = o mu (1 - d2)^beta,
//           skipped when d2 < 0, d2 >= 1 or |alpha| < 1e-5.
template <typename S, bool BETA>
__device__ __forceinline__ bool eval_profile(const Prim2D<S>& p, S px, S py, S& vis, S& alpha) {
    if (!BETA) {
        const S dx = px - p.mx, dy = py - p.my;
        const S sigma = S(0.5) * (p.a * dx * dx + p.c * dy * dy) + p.b * dx * dy;
        vis = Math<S>::exp(-sigma);
        alpha = p.opac * p.mu * vis;
        return !(sigma < S(0) || fabs(alpha) < S(1e-5));
    } else {
        const S dx = p.mx - px, dy = p.my - py;
        const S d2 = p.a * dx * dx + S(2) * p.b * dx * dy + p.c * dy * dy;
        if (d2 < S(0) || d2 >= S(1)) return false;
        const S h = S(1) - d2;
        vis = Math<S>::pow(h, p.beta);
        alpha = p.opac * p.mu * vis;
        return !(fabs(alpha) < S(1e-5));
    }
}

template <typename S>
struct Grad2D {
    S xy[2] = {S(0), S(0)};
    S abs[2] = {S(0), S(0)};
    S conic[3] = {S(0), S(0), S(0)};
    S opac = S(0), mu = S(0), beta = S(0);

    __device__ __forceinline__ void reduce_wave(bool with_abs) {
        xy[0] = wave_sum(xy[0]);
        xy[1] = wave_sum(xy[1]);
        if (with_abs) {
            abs[0] = wave_sum(abs[0]);
            abs[1] = wave_sum(abs[1]);
        }
        conic[0] = wave_sum(conic[0]);
        conic[1] = wave_sum(conic[1]);
        conic[2] = wave_sum(conic[2]);
        opac = wave_sum(opac);
        mu = wave_sum(mu);
        beta = wave_sum(beta);
    }

    __device__ __forceinline__ void commit(int32_t g, S* v_abs, S* v_xy, S* v_conic, S* v_opac,
                                           S* v_mu, S* v_beta) const {
        const size_t G = size_t(g);
        atomic_add(v_xy + 2 * G, xy[0]);
        atomic_add(v_xy + 2 * G + 1, xy[1]);
        if (v_abs) {
            atomic_add(v_abs + 2 * G, abs[0]);
            atomic_add(v_abs + 2 * G + 1, abs[1]);
        }
        atomic_add(v_conic + 3 * G, conic[0]);
        atomic_add(v_conic + 3 * G + 1, conic[1]);
        atomic_add(v_conic + 3 * G + 2, conic[2]);
        atomic_add(v_opac + G, opac);
        atomic_add(v_mu + G, mu);
        if (v_beta) atomic_add(v_beta + G, beta);
    }
};

// Adds one pixel's VJP for upstream adjoint v (alpha, vis from eval_profile).
template <typename S, bool BETA>
__device__ __forceinline__ void pixel_vjp(const Prim2D<S>& p, S px, S py, S vis, S alpha, S v,
                                          Grad2D<S>& g) {
    S gx, gy;
    if (!BETA) {
        const S dx = px - p.mx, dy = py - p.my;
        const S vs = -v * alpha;  // d/d sigma
        gx = -vs * (p.a * dx + p.b * dy);
        gy = -vs * (p.c * dy + p.b * dx);
        g.conic[0] += vs * S(0.5) * dx * dx;
        g.conic[1] += vs * dx * dy;
        g.conic[2] += vs * S(0.5) * dy * dy;
    } else {
        const S dx = p.mx - px, dy = p.my - py;
        const S d2 = p.a * dx * dx + S(2) * p.b * dx * dy + p.c * dy * dy;
        const S h = S(1) - d2;
        const S vd = -p.opac * p.mu * p.beta * (vis / h) * v;  // d/d d2
        gx = S(2) * vd * (p.a * dx + p.b * dy);
        gy = S(2) * vd * (p.b * dx + p.c * dy);
        g.conic[0] += vd * dx * dx;
        g.conic[1] += S(2) * vd * dx * dy;
        g.conic[2] += vd * dy * dy;
        g.beta += v * alpha * Math<S>::log(h);
    }
    g.xy[0] += gx;
    g.xy[1] += gy;
    g.abs[0] += fabs(gx);
    g.abs[1] += fabs(gy);
    g.opac += v * p.mu * vis;
    g.mu += v * p.opac * vis;
}

}  // namespace r2s

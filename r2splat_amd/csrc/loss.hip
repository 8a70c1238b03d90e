// Volume TV loss (anisotropic, edge-mean normalised; isotropic, voxel-mean
// normalised) and mean-L1 image loss.
// Semantics: r2splat/cuda/csrc/VolumeLossFwd.cu:39-153, VolumeLossBwd.cu:46-115,
// Loss.cpp:115-188, ImageLossFwd.cu:20-31, ImageLossBwd.cu:14-30.
#include "common.hpp"
#include "raster_common.hpp"

namespace r2s {

constexpr int kLossThreads = 256;

template <typename S>
__device__ __forceinline__ S sgn(S x) { return x > S(0) ? S(1) : (x < S(0) ? S(-1) : S(0)); }

// Block sum of one value per thread, then one atomic per block.
template <typename S>
__device__ __forceinline__ void block_sum_atomic(S v, S* out) {
    __shared__ S part[kLossThreads / kWave];
    v = wave_sum(v);
    if (lane_id() == 0) part[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        S t = S(0);
        for (int w = 0; w < kLossThreads / kWave; ++w) t += part[w];
        atomic_add(out, t);
    }
}

template <typename S>
__global__ void __launch_bounds__(kLossThreads)
tv_fwd_kernel(const S* __restrict__ vol, uint32_t B, uint32_t D, uint32_t H, uint32_t W,
              S* __restrict__ tv_map, S* __restrict__ tv_sum) {
    const size_t total = size_t(B) * D * H * W;
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    S tv = S(0);
    if (idx < total) {
        const uint32_t w = uint32_t(idx % W), h = uint32_t((idx / W) % H), d = uint32_t((idx / (size_t(W) * H)) % D);
        const S v = vol[idx];
        if (d + 1 < D) tv += fabs(vol[idx + size_t(H) * W] - v);
        if (h + 1 < H) tv += fabs(vol[idx + W] - v);
        if (w + 1 < W) tv += fabs(vol[idx + 1] - v);
        tv_map[idx] = tv;
    }
    block_sum_atomic(tv, tv_sum);
}

// grad = ((sign sum) * (g / edges)) * lambda: the reference's rounding order
// (the kernel's g / edges product, then the host-side multiply by lambda).
template <typename S>
__global__ void __launch_bounds__(kLossThreads)
tv_bwd_kernel(const S* __restrict__ vol, const S* __restrict__ grad_out, uint32_t B, uint32_t D, uint32_t H,
              uint32_t W, S edges, S lambda, S* __restrict__ grad) {
    const size_t total = size_t(B) * D * H * W;
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint32_t w = uint32_t(idx % W), h = uint32_t((idx / W) % H), d = uint32_t((idx / (size_t(W) * H)) % D);
    const S v = vol[idx];
    const size_t sd = size_t(H) * W, sh = W;
    S g = S(0);
    if (d + 1 < D) g -= sgn(vol[idx + sd] - v);
    if (d > 0) g += sgn(v - vol[idx - sd]);
    if (h + 1 < H) g -= sgn(vol[idx + sh] - v);
    if (h > 0) g += sgn(v - vol[idx - sh]);
    if (w + 1 < W) g -= sgn(vol[idx + 1] - v);
    if (w > 0) g += sgn(v - vol[idx - 1]);
    const S per_edge = grad_out[0] / edges;
    grad[idx] = (g * per_edge) * lambda;
}

// Isotropic TV: t(v) = sqrt(dz^2 + dy^2 + dx^2 + eps) with forward differences
// (zero past the last index of an axis); tv_map = t, tv_sum = sum of t.
template <typename S>
__device__ __forceinline__ S iso_term(const S* __restrict__ vol, size_t idx, uint32_t d, uint32_t h, uint32_t w,
                                      uint32_t D, uint32_t H, uint32_t W, S eps, S& dz, S& dy, S& dx) {
    const S v = vol[idx];
    dz = d + 1 < D ? vol[idx + size_t(H) * W] - v : S(0);
    dy = h + 1 < H ? vol[idx + W] - v : S(0);
    dx = w + 1 < W ? vol[idx + 1] - v : S(0);
    return sqrt(dz * dz + dy * dy + dx * dx + eps);
}

template <typename S>
__global__ void __launch_bounds__(kLossThreads)
tv_iso_fwd_kernel(const S* __restrict__ vol, uint32_t B, uint32_t D, uint32_t H, uint32_t W, S eps,
                  S* __restrict__ tv_map, S* __restrict__ tv_sum) {
    const size_t total = size_t(B) * D * H * W;
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    S t = S(0);
    if (idx < total) {
        const uint32_t w = uint32_t(idx % W), h = uint32_t((idx / W) % H), d = uint32_t((idx / (size_t(W) * H)) % D);
        S dz, dy, dx;
        t = iso_term(vol, idx, d, h, w, D, H, W, eps, dz, dy, dx);
        tv_map[idx] = t;
    }
    block_sum_atomic(t, tv_sum);
}

// d/dV(v) of sum_u t(u): -(dz + dy + dx)(v) / t(v) plus dz(v - ez) / t(v - ez),
// dy(v - ey) / t(v - ey) and dx(v - ex) / t(v - ex) where those voxels exist.
template <typename S>
__global__ void __launch_bounds__(kLossThreads)
tv_iso_bwd_kernel(const S* __restrict__ vol, const S* __restrict__ grad_out, uint32_t B, uint32_t D, uint32_t H,
                  uint32_t W, S eps, S scale, S* __restrict__ grad) {
    const size_t total = size_t(B) * D * H * W;
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const uint32_t w = uint32_t(idx % W), h = uint32_t((idx / W) % H), d = uint32_t((idx / (size_t(W) * H)) % D);
    const size_t sd = size_t(H) * W, sh = W;
    S dz, dy, dx;
    S t = iso_term(vol, idx, d, h, w, D, H, W, eps, dz, dy, dx);
    S g = -(dz + dy + dx) / t;
    if (d > 0) {
        t = iso_term(vol, idx - sd, d - 1, h, w, D, H, W, eps, dz, dy, dx);
        g += dz / t;
    }
    if (h > 0) {
        t = iso_term(vol, idx - sh, d, h - 1, w, D, H, W, eps, dz, dy, dx);
        g += dy / t;
    }
    if (w > 0) {
        t = iso_term(vol, idx - 1, d, h, w - 1, D, H, W, eps, dz, dy, dx);
        g += dx / t;
    }
    grad[idx] = g * (grad_out[0] * scale);
}

template <typename S>
__global__ void __launch_bounds__(kLossThreads)
l1_fwd_kernel(const S* __restrict__ p, const S* __restrict__ t, uint64_t n, S* __restrict__ map,
              S* __restrict__ sum) {
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    S a = S(0);
    if (idx < n) {
        a = fabs(p[idx] - t[idx]);
        map[idx] = a;
    }
    block_sum_atomic(a, sum);
}

template <typename S>
__global__ void __launch_bounds__(kLossThreads)
l1_bwd_kernel(const S* __restrict__ p, const S* __restrict__ t, const S* __restrict__ g, uint64_t n,
              S lambda, S* __restrict__ out) {
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    out[idx] = g[0] * lambda * sgn(p[idx] - t[idx]) / S(n);  // the reference's operation order
}

}  // namespace r2s

using namespace r2s;

extern "C" int r2s_volume_tv_fwd(int dtype, const void* volume, uint32_t B, uint32_t D, uint32_t H,
                                 uint32_t W, void* tv_map, void* tv_sum, r2s_stream_t stream) {
    const size_t total = size_t(B) * D * H * W;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (int rc = fill_zero<S>((S*)tv_sum, 1, stream)) return rc;
        if (total == 0) return R2S_OK;
        tv_fwd_kernel<S><<<ceil_div(total, kLossThreads), kLossThreads, 0, stream>>>(
            (const S*)volume, B, D, H, W, (S*)tv_map, (S*)tv_sum);
        return check_launch("volume_tv_fwd");
    });
}

extern "C" int r2s_volume_tv_bwd(int dtype, const void* volume, const void* grad_output, uint32_t B,
                                 uint32_t D, uint32_t H, uint32_t W, float lambda_tv,
                                 void* grad_volume, r2s_stream_t stream) {
    const size_t total = size_t(B) * D * H * W;
    if (total == 0) return R2S_OK;
    const int64_t edges64 = int64_t(B) * ((int64_t(D) - 1) * H * W + int64_t(D) * (H - 1) * W +
                                          int64_t(D) * H * (W - 1));
    const int64_t edges = edges64 < 1 ? 1 : edges64;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        tv_bwd_kernel<S><<<ceil_div(total, kLossThreads), kLossThreads, 0, stream>>>(
            (const S*)volume, (const S*)grad_output, B, D, H, W, S(edges), S(lambda_tv), (S*)grad_volume);
        return check_launch("volume_tv_bwd");
    });
}

extern "C" int r2s_volume_tv_iso_fwd(int dtype, const void* volume, uint32_t B, uint32_t D, uint32_t H,
                                     uint32_t W, double eps, void* tv_map, void* tv_sum, r2s_stream_t stream) {
    R2S_REQUIRE(eps > 0.0, "isotropic TV needs eps > 0 (got %g)", eps);
    const size_t total = size_t(B) * D * H * W;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (int rc = fill_zero<S>((S*)tv_sum, 1, stream)) return rc;
        if (total == 0) return R2S_OK;
        tv_iso_fwd_kernel<S><<<ceil_div(total, kLossThreads), kLossThreads, 0, stream>>>(
            (const S*)volume, B, D, H, W, S(eps), (S*)tv_map, (S*)tv_sum);
        return check_launch("volume_tv_iso_fwd");
    });
}

extern "C" int r2s_volume_tv_iso_bwd(int dtype, const void* volume, const void* grad_output, uint32_t B,
                                     uint32_t D, uint32_t H, uint32_t W, double eps, double lambda_tv,
                                     void* grad_volume, r2s_stream_t stream) {
    R2S_REQUIRE(eps > 0.0, "isotropic TV needs eps > 0 (got %g)", eps);
    const size_t total = size_t(B) * D * H * W;
    if (total == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        tv_iso_bwd_kernel<S><<<ceil_div(total, kLossThreads), kLossThreads, 0, stream>>>(
            (const S*)volume, (const S*)grad_output, B, D, H, W, S(eps), S(lambda_tv / double(total)),
            (S*)grad_volume);
        return check_launch("volume_tv_iso_bwd");
    });
}

extern "C" int r2s_image_l1_fwd(int dtype, const void* prediction, const void* target, uint64_t n,
                                void* l1_map, void* l1_sum, r2s_stream_t stream) {
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (int rc = fill_zero<S>((S*)l1_sum, 1, stream)) return rc;
        if (n == 0) return R2S_OK;
        l1_fwd_kernel<S><<<ceil_div(n, kLossThreads), kLossThreads, 0, stream>>>(
            (const S*)prediction, (const S*)target, n, (S*)l1_map, (S*)l1_sum);
        return check_launch("image_l1_fwd");
    });
}

extern "C" int r2s_image_l1_bwd(int dtype, const void* prediction, const void* target,
                                const void* grad_output, uint64_t n, float lambda_l1,
                                void* grad_prediction, r2s_stream_t stream) {
    if (n == 0) return R2S_OK;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        l1_bwd_kernel<S><<<ceil_div(n, kLossThreads), kLossThreads, 0, stream>>>(
            (const S*)prediction, (const S*)target, (const S*)grad_output, n, S(lambda_l1),
            (S*)grad_prediction);
        return check_launch("image_l1_bwd");
    });
}

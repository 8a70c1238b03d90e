// Training-step kernels for graph-captured fits: the MSE loss and its adjoint,
// and an Adam update over a flat parameter buffer whose step counter lives on
// the device (so a captured graph can be replayed without host-side scalars).
//
// The reference leaves losses on images and the optimizer to the caller
// (r2splat/explicit.py:1-20); these kernels exist for the batched fits.
#include "common.hpp"
#include "raster_common.hpp"

namespace r2s {

constexpr int kTrainThreads = 256;
constexpr uint32_t kMseItems = 8;  // elements per thread per block pass
constexpr uint32_t kMseBlockElems = kTrainThreads * kMseItems;

// v = scale (img - tgt); partial[b] = sum over block b of (img - tgt)^2 (fixed order).
__global__ void __launch_bounds__(kTrainThreads)
mse_adjoint_kernel(const float* __restrict__ img, const float* __restrict__ tgt, uint64_t n, float scale,
                   float* __restrict__ v, float* __restrict__ partial) {
    __shared__ float part[kTrainThreads / kWave];
    const uint64_t base = uint64_t(blockIdx.x) * kMseBlockElems;
    float s = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < kMseItems; ++k) {
        const uint64_t i = base + uint64_t(k) * kTrainThreads + threadIdx.x;
        if (i < n) {
            const float d = img[i] - tgt[i];
            v[i] = scale * d;
            s = fmaf(d, d, s);
        }
    }
    if (partial) {
        s = wave_sum(s);
        if (lane_id() == 0) part[threadIdx.x / kWave] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            float t = 0.f;
            for (int w = 0; w < kTrainThreads / kWave; ++w) t += part[w];
            partial[blockIdx.x] = t;
        }
    }
}

// out[0] = inv_n * sum(partial[0..m)), one block, fixed order (double accumulation).
__global__ void __launch_bounds__(kTrainThreads)
sum_partials_kernel(const float* __restrict__ partial, uint32_t m, float inv_n, float* __restrict__ out) {
    __shared__ double part[kTrainThreads];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < m; i += kTrainThreads) s += double(partial[i]);
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kTrainThreads; ++k) t += part[k];
        out[0] = float(t * double(inv_n));
    }
}

// Adam (bias-corrected, as torch.optim.Adam without weight decay). The step
// about to be taken is state[0] + 1; when `skip` is set (the forward of this
// step overflowed its intersection capacity) nothing is updated.
__global__ void __launch_bounds__(kTrainThreads)
adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
            uint64_t n, float lr, float b1, float b2, float eps, const uint32_t* __restrict__ state,
            const uint32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = float(state[0] + 1u);
    const float bc1 = 1.f - powf(b1, t), bc2 = 1.f - powf(b2, t);
    const float gi = g[i];
    const float mi = fmaf(b1, m[i], (1.f - b1) * gi);
    const float vi = fmaf(b2, v[i], (1.f - b2) * gi * gi);
    m[i] = mi;
    v[i] = vi;
    const float denom = sqrtf(vi) / sqrtf(bc2) + eps;
    p[i] -= (lr / bc1) * mi / denom;
}

// state[0] counts applied steps, state[1] skipped ones.
__global__ void adam_finish_kernel(uint32_t* __restrict__ state, const uint32_t* __restrict__ skip) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (skip && *skip) state[1] += 1u;
        else state[0] += 1u;
    }
}

__global__ void __launch_bounds__(kTrainThreads)
clamp_min_kernel(float* __restrict__ x, uint64_t n, float lo) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = fmaxf(x[i], lo);
}

}  // namespace r2s

using namespace r2s;

extern "C" int r2s_clamp_min(float* x, uint64_t n, float lo, r2s_stream_t stream) {
    if (n == 0) return R2S_OK;
    R2S_REQUIRE(n / kTrainThreads < (1u << 31), "buffer too large");
    clamp_min_kernel<<<ceil_div(n, kTrainThreads), kTrainThreads, 0, stream>>>(x, n, lo);
    return check_launch("clamp_min");
}

extern "C" size_t r2s_mse_partials(uint64_t n) { return (n + kMseBlockElems - 1) / kMseBlockElems; }

extern "C" int r2s_mse_adjoint(const float* image, const float* target, uint64_t n, float* v_image,
                               float* partials, float* loss, r2s_stream_t stream) {
    R2S_REQUIRE((partials == nullptr) == (loss == nullptr), "partials and loss go together");
    if (n == 0) return loss ? fill_zero<float>(loss, 1, stream) : R2S_OK;
    const uint64_t blocks = (n + kMseBlockElems - 1) / kMseBlockElems;
    R2S_REQUIRE(blocks < (1u << 31), "image too large");
    mse_adjoint_kernel<<<uint32_t(blocks), kTrainThreads, 0, stream>>>(image, target, n, float(2.0 / double(n)),
                                                                       v_image, partials);
    if (int rc = check_launch("mse_adjoint")) return rc;
    if (loss) {
        sum_partials_kernel<<<1, kTrainThreads, 0, stream>>>(partials, uint32_t(blocks), float(1.0 / double(n)),
                                                             loss);
        return check_launch("mse_sum");
    }
    return R2S_OK;
}

extern "C" int r2s_adam_step(float* params, const float* grads, float* exp_avg, float* exp_avg_sq, uint64_t n,
                             float lr, float beta1, float beta2, float eps, uint32_t* state,
                             const uint32_t* skip_flag, r2s_stream_t stream) {
    R2S_REQUIRE(state != nullptr, "adam state counter is required");
    R2S_REQUIRE(n / kTrainThreads < (1u << 31), "parameter buffer too large");
    if (n > 0) {
        adam_kernel<<<ceil_div(n, kTrainThreads), kTrainThreads, 0, stream>>>(params, grads, exp_avg, exp_avg_sq, n,
                                                                              lr, beta1, beta2, eps, state,
                                                                              skip_flag);
        if (int rc = check_launch("adam")) return rc;
    }
    adam_finish_kernel<<<1, kWave, 0, stream>>>(state, skip_flag);
    return check_launch("adam_finish");
}

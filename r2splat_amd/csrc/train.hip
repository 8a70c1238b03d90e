// Training-step kernels for graph-captured fits: the MSE loss and its adjoint,
// and an Adam update over a flat parameter buffer whose step counter lives on
// the device (so a captured graph can be replayed without host-side scalars).
//
// The reference leaves losses on images and the optimizer to the caller
// (r2splat/explicit.py:1-20); these kernels exist for the batched fits.
#include "common.hpp"
#include "raster_common.hpp"

#include <atomic>
#include <mutex>

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

// Loss partials of r2s_mse_adjoint_noalloc: one file-scope array per GPU (the
// code object's own device memory, held until the process ends) in place of
// a caller workspace. 2^16 partials of 2048 pixels each cover images of up to
// 134,217,728 pixels in 256 KiB.
constexpr uint32_t kStaticMsePartials = 1u << 16;
static __device__ __align__(256) float scratch[kStaticMsePartials];

// Device address of `scratch` on each GPU. In host code the array's name is a
// host-side stand-in, never a device address: kernels get these pointers.
constexpr int kMaxDevices = 64;
static std::atomic<float*> g_scratch[kMaxDevices];
static std::mutex g_scratch_lock;

// The device that kernels launched on `stream` run on.
static int stream_device(r2s_stream_t stream, int* device) {
    if (stream == nullptr) {
        R2S_HIP(hipGetDevice(device));
        return R2S_OK;
    }
    hipDevice_t d = 0;
    R2S_HIP(hipStreamGetDevice(stream, &d));
    *device = int(d);
    return R2S_OK;
}

// Looks up (once per device, with that device current) the device address of
// `scratch` on the device `stream` runs on.
static int static_partials(r2s_stream_t stream, float** out) {
    int device = 0;
    if (int rc = stream_device(stream, &device)) return rc;
    R2S_REQUIRE(device >= 0 && device < kMaxDevices, "device ordinal %d beyond the %d supported", device,
                kMaxDevices);
    float* p = g_scratch[device].load(std::memory_order_acquire);
    if (p == nullptr) {
        std::lock_guard<std::mutex> guard(g_scratch_lock);
        p = g_scratch[device].load(std::memory_order_relaxed);
        if (p == nullptr) {
            int current = 0;
            R2S_HIP(hipGetDevice(&current));
            if (current != device) R2S_HIP(hipSetDevice(device));
            void* addr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
            if (current != device) R2S_HIP(hipSetDevice(current));
            if (e != hipSuccess || addr == nullptr) {
                set_error("mse_adjoint_noalloc: hipGetSymbolAddress failed on device %d: %s", device,
                          hipGetErrorString(e));
                return R2S_ERR_HIP;
            }
            p = static_cast<float*>(addr);
            g_scratch[device].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return R2S_OK;
}

extern "C" int r2s_mse_adjoint_noalloc(const float* image, const float* target, uint64_t n, float* v_image,
                                       float* loss, r2s_stream_t stream) {
    R2S_REQUIRE(loss != nullptr, "mse_adjoint_noalloc: loss is required");
    const uint64_t limit = uint64_t(kStaticMsePartials) * kMseBlockElems;
    R2S_REQUIRE(n <= limit, "mse_adjoint_noalloc: %llu pixels exceed the static partial buffer's %llu",
                (unsigned long long)n, (unsigned long long)limit);
    if (n == 0) return fill_zero<float>(loss, 1, stream);
    float* partials = nullptr;
    if (int rc = static_partials(stream, &partials)) return rc;
    return r2s_mse_adjoint(image, target, n, v_image, partials, loss, stream);
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

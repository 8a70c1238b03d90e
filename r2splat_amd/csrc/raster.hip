// 2D CT rasterization: each pixel sums opacity * mu * profile over the
// primitives binned to its tile (additive line integrals, no compositing).
//
// Operator-level kernels for every tile size and both float/double:
//  * tile-parallel forward (mean, mean+variance), staging the tile's primitive
//    list through LDS in batches;
//  * tile-parallel ("pixel" mode) VJP and the uncertainty VJP: one wave64
//    reduction per (wave, primitive) followed by one atomic per field;
//  * intersection-parallel ("primitive" mode) VJP: one wave64 per
//    intersection, lanes stride the tile's pixels, one atomic per field.
// Profiles follow r2splat/cuda/csrc/RasterizeToPixelsFwd.cu:155-180 (Gaussian)
// and :664-680 (Beta), and the VJPs RasterizeToPixelsBwd.cu:169-265,421-501,
// 674-706,909-945.
#include "common.hpp"
#include "raster_common.hpp"

namespace r2s {

constexpr uint32_t kBatch = 256;

template <typename S>
struct PrimLDS {
    S* mx; S* my; S* a; S* b; S* c; S* opac; S* mu; S* beta; S* q; int32_t* id;
};

template <typename S>
__device__ __forceinline__ PrimLDS<S> carve_lds(unsigned char* smem, uint32_t batch) {
    PrimLDS<S> p;
    S* f = reinterpret_cast<S*>(smem);
    p.mx = f; p.my = f + batch; p.a = f + 2 * batch; p.b = f + 3 * batch; p.c = f + 4 * batch;
    p.opac = f + 5 * batch; p.mu = f + 6 * batch; p.beta = f + 7 * batch; p.q = f + 8 * batch;
    p.id = reinterpret_cast<int32_t*>(f + 9 * batch);
    return p;
}

inline size_t lds_bytes(int dtype, uint32_t batch) {
    const size_t s = dtype == R2S_F64 ? 8 : 4;
    return batch * (9 * s + 4);
}

template <typename S>
struct RasterArgs {
    const S* means2d; const S* conics; const S* opac; const S* mus; const S* betas; const S* unc;
    const int32_t* offsets; const int32_t* flatten; uint32_t n_isects, C, N, W, H, T, ntx, nty;
};

// Decodes block -> (camera, tile) and returns the tile's [start, end) range.
template <typename S>
__device__ __forceinline__ void tile_range(const RasterArgs<S>& A, uint32_t blk, uint32_t& cam,
                                           uint32_t& ty, uint32_t& tx, int32_t& start, int32_t& end) {
    const uint32_t nt = A.ntx * A.nty;
    cam = blk / nt;
    const uint32_t tile = blk - cam * nt;
    ty = tile / A.ntx;
    tx = tile - ty * A.ntx;
    const uint64_t key = uint64_t(cam) * nt + tile;
    start = A.offsets[key];
    end = (key + 1 == uint64_t(A.C) * nt) ? int32_t(A.n_isects) : A.offsets[key + 1];
    start = max(0, min(start, int32_t(A.n_isects)));
    end = max(start, min(end, int32_t(A.n_isects)));
}

template <typename S, bool BETA, bool VAR>
__device__ __forceinline__ void stage_batch(const RasterArgs<S>& A, const PrimLDS<S>& L, int32_t b0,
                                            int32_t end, uint32_t batch) {
    const uint32_t t = threadIdx.x;
    if (t < batch && b0 + int32_t(t) < end) {
        int32_t g = A.flatten[b0 + t];
        const bool ok = g >= 0 && uint32_t(g) < A.C * A.N;
        if (!ok) g = 0;  // out-of-range ids contribute nothing
        L.id[t] = g;
        L.mx[t] = A.means2d[2 * size_t(g)];
        L.my[t] = A.means2d[2 * size_t(g) + 1];
        L.a[t] = A.conics[3 * size_t(g)];
        L.b[t] = A.conics[3 * size_t(g) + 1];
        L.c[t] = A.conics[3 * size_t(g) + 2];
        L.opac[t] = ok ? A.opac[g] : S(0);
        L.mu[t] = A.mus[g];
        if (BETA) L.beta[t] = A.betas[g];
        if (VAR) L.q[t] = A.unc[uint32_t(g) % A.N];
    }
}

template <typename S, bool BETA>
__device__ __forceinline__ Prim2D<S> lds_prim(const PrimLDS<S>& L, uint32_t t) {
    Prim2D<S> p;
    p.mx = L.mx[t]; p.my = L.my[t]; p.a = L.a[t]; p.b = L.b[t]; p.c = L.c[t];
    p.opac = L.opac[t]; p.mu = L.mu[t]; p.beta = BETA ? L.beta[t] : S(0);
    return p;
}

// ---------------------------------------------------------------------------
// Forward (mean, optional variance). Block = one tile, T*T threads.
// ---------------------------------------------------------------------------
template <typename S, bool BETA, bool VAR>
__global__ void raster_fwd_kernel(RasterArgs<S> A, S* __restrict__ out, S* __restrict__ out_var) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t batch = min(blockDim.x, kBatch);
    PrimLDS<S> L = carve_lds<S>(smem, batch);
    uint32_t cam, ty, tx;
    int32_t start, end;
    tile_range(A, blockIdx.x, cam, ty, tx, start, end);
    const uint32_t i = ty * A.T + threadIdx.x / A.T;
    const uint32_t j = tx * A.T + threadIdx.x % A.T;
    const bool inside = (threadIdx.x < A.T * A.T) && i < A.H && j < A.W;
    const S px = S(j) + S(0.5), py = S(i) + S(0.5);
    S acc = S(0), acc_var = S(0);
    for (int32_t b0 = start; b0 < end; b0 += int32_t(batch)) {
        __syncthreads();
        stage_batch<S, BETA, VAR>(A, L, b0, end, batch);
        __syncthreads();
        if (!inside) continue;
        const uint32_t cnt = uint32_t(min<int32_t>(int32_t(batch), end - b0));
        for (uint32_t t = 0; t < cnt; ++t) {
            const Prim2D<S> p = lds_prim<S, BETA>(L, t);
            S vis, alpha;
            if (!eval_profile<S, BETA>(p, px, py, vis, alpha)) continue;
            acc += alpha;
            if (VAR) acc_var += L.q[t] * alpha * alpha;
        }
    }
    if (inside) {
        const size_t o = (size_t(cam) * A.H + i) * A.W + j;
        out[o] = acc;
        if (VAR) out_var[o] = acc_var;
    }
}

// ---------------------------------------------------------------------------
// Tile-parallel VJPs. MODE 0: mean VJP (pixel mode); MODE 1: uncertainty VJP.
// ---------------------------------------------------------------------------
template <typename S, bool BETA, int MODE>
__global__ void raster_bwd_tile_kernel(RasterArgs<S> A, const S* __restrict__ v_render,
                                       S* __restrict__ v_abs, S* __restrict__ v_xy,
                                       S* __restrict__ v_conic, S* __restrict__ v_opac,
                                       S* __restrict__ v_mu, S* __restrict__ v_beta,
                                       S* __restrict__ v_q) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t batch = min(blockDim.x, kBatch);
    PrimLDS<S> L = carve_lds<S>(smem, batch);
    uint32_t cam, ty, tx;
    int32_t start, end;
    tile_range(A, blockIdx.x, cam, ty, tx, start, end);
    const uint32_t i = ty * A.T + threadIdx.x / A.T;
    const uint32_t j = tx * A.T + threadIdx.x % A.T;
    const bool inside = (threadIdx.x < A.T * A.T) && i < A.H && j < A.W;
    const S px = S(j) + S(0.5), py = S(i) + S(0.5);
    const S v = inside ? v_render[(size_t(cam) * A.H + i) * A.W + j] : S(0);
    const bool lane0 = lane_id() == 0;
    for (int32_t b0 = start; b0 < end; b0 += int32_t(batch)) {
        __syncthreads();
        stage_batch<S, BETA, false>(A, L, b0, end, batch);
        __syncthreads();
        const uint32_t cnt = uint32_t(min<int32_t>(int32_t(batch), end - b0));
        for (uint32_t t = 0; t < cnt; ++t) {
            const Prim2D<S> p = lds_prim<S, BETA>(L, t);
            S vis = S(0), alpha = S(0);
            const bool valid = inside && eval_profile<S, BETA>(p, px, py, vis, alpha);
            const int32_t g = L.id[t];
            if (MODE == 1) {
                const S s = wave_sum(valid ? v * alpha * alpha : S(0));
                if (lane0) atomic_add(v_q + uint32_t(g) % A.N, s);
                continue;
            }
            Grad2D<S> gr;
            if (valid) pixel_vjp<S, BETA>(p, px, py, vis, alpha, v, gr);
            gr.reduce_wave(v_abs != nullptr);
            if (lane0) gr.commit(g, v_abs, v_xy, v_conic, v_opac, v_mu, v_beta);
        }
    }
}

// ---------------------------------------------------------------------------
// Intersection-parallel VJP: one wave64 per intersection.
// ---------------------------------------------------------------------------
template <typename S, bool BETA>
__global__ void __launch_bounds__(256)
raster_bwd_isect_kernel(RasterArgs<S> A, const int64_t* __restrict__ isect_ids,
                        const S* __restrict__ v_render, const S* __restrict__ v_var,
                        const S* __restrict__ unc, S* __restrict__ v_abs, S* __restrict__ v_xy,
                        S* __restrict__ v_conic, S* __restrict__ v_opac, S* __restrict__ v_mu,
                        S* __restrict__ v_beta, S* __restrict__ v_q) {
    // grid-stride over intersections (one wave each): a launch covers at most
    // kIsectBlocks blocks, whatever the intersection count
    const uint32_t waves = blockDim.x / kWave;
    const uint64_t stride = uint64_t(gridDim.x) * waves;
    for (uint64_t e = uint64_t(blockIdx.x) * waves + threadIdx.x / kWave; e < A.n_isects; e += stride) {
        const uint32_t nt = A.ntx * A.nty;
        const int64_t key = isect_ids[e];
        const uint32_t cam = uint32_t(key / nt);
        if (cam >= A.C) continue;
        const uint32_t tile = uint32_t(key - int64_t(cam) * nt);
        const uint32_t ty = tile / A.ntx, tx = tile - (tile / A.ntx) * A.ntx;
        const int32_t g = A.flatten[e];
        if (g < 0 || uint32_t(g) >= A.C * A.N) continue;
        Prim2D<S> p;
        p.mx = A.means2d[2 * size_t(g)];
        p.my = A.means2d[2 * size_t(g) + 1];
        p.a = A.conics[3 * size_t(g)];
        p.b = A.conics[3 * size_t(g) + 1];
        p.c = A.conics[3 * size_t(g) + 2];
        p.opac = A.opac[g];
        p.mu = A.mus[g];
        p.beta = BETA ? A.betas[g] : S(0);
        const uint32_t n = uint32_t(g) % A.N;
        const S q = unc ? unc[n] : S(0);
        const size_t img = size_t(cam) * A.H * A.W;
        Grad2D<S> gr;
        S vq = S(0);
        const uint32_t npix = A.T * A.T;
        for (uint32_t k = lane_id(); k < npix; k += kWave) {
            const uint32_t ly = k / A.T, lx = k - ly * A.T;
            const uint32_t i = ty * A.T + ly, j = tx * A.T + lx;
            if (i >= A.H || j >= A.W) continue;
            const S px = S(j) + S(0.5), py = S(i) + S(0.5);
            S vis, alpha;
            if (!eval_profile<S, BETA>(p, px, py, vis, alpha)) continue;
            const size_t o = img + size_t(i) * A.W + j;
            S v = v_render[o];
            if (v_var) {
                const S vv = v_var[o];
                vq += vv * alpha * alpha;
                if (unc) v += S(2) * q * alpha * vv;
            }
            pixel_vjp<S, BETA>(p, px, py, vis, alpha, v, gr);
        }
        gr.reduce_wave(v_abs != nullptr);
        if (v_q) vq = wave_sum(vq);
        if (lane_id() == 0) {
            gr.commit(g, v_abs, v_xy, v_conic, v_opac, v_mu, v_beta);
            if (v_q) atomic_add(v_q + n, vq);
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
inline int check_tile(uint32_t T, uint32_t ntx, uint32_t nty, uint32_t W, uint32_t H) {
    R2S_REQUIRE(T >= 1 && T * T <= 1024, "tile_size %u must satisfy 1 <= tile_size^2 <= 1024", T);
    R2S_REQUIRE(uint64_t(ntx) * T >= W && uint64_t(nty) * T >= H,
                "tile grid %ux%u of size %u does not cover the %ux%u image", ntx, nty, T, W, H);
    return R2S_OK;
}

inline uint32_t block_threads(uint32_t T) { return ((T * T + kWave - 1) / kWave) * kWave; }

template <typename S>
RasterArgs<S> make_args(const void* means2d, const void* conics, const void* opac, const void* mus,
                        const void* betas, const void* unc, const int32_t* offsets,
                        const int32_t* flatten, uint32_t n_isects, uint32_t C, uint32_t N,
                        uint32_t W, uint32_t H, uint32_t T, uint32_t ntx, uint32_t nty) {
    return RasterArgs<S>{(const S*)means2d, (const S*)conics, (const S*)opac, (const S*)mus,
                         (const S*)betas, (const S*)unc, offsets, flatten, n_isects, C, N, W, H, T,
                         ntx, nty};
}

int launch_fwd(int dtype, bool beta, bool var, const void* means2d, const void* conics,
               const void* opac, const void* mus, const void* betas, const void* unc,
               const int32_t* offsets, const int32_t* flatten, uint32_t n_isects, uint32_t C,
               uint32_t N, uint32_t W, uint32_t H, uint32_t T, uint32_t ntx, uint32_t nty,
               void* out, void* out_var, hipStream_t stream) {
    if (int rc = check_tile(T, ntx, nty, W, H)) return rc;
    const uint64_t blocks = uint64_t(C) * ntx * nty;
    if (blocks == 0 || W == 0 || H == 0) return R2S_OK;
    const uint32_t threads = block_threads(T);
    const size_t shm = lds_bytes(dtype, std::min(threads, kBatch));
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        auto A = make_args<S>(means2d, conics, opac, mus, betas, unc, offsets, flatten, n_isects, C,
                              N, W, H, T, ntx, nty);
        auto k = beta ? (var ? raster_fwd_kernel<S, true, true> : raster_fwd_kernel<S, true, false>)
                      : (var ? raster_fwd_kernel<S, false, true> : raster_fwd_kernel<S, false, false>);
        k<<<uint32_t(blocks), threads, shm, stream>>>(A, (S*)out, (S*)out_var);
        return check_launch("rasterize_to_pixels_fwd");
    });
}

}  // namespace r2s

using namespace r2s;

extern "C" int r2s_rasterize_to_pixels_fwd(int dtype, const void* means2d, const void* conics,
                                           const void* opacities, const void* mus,
                                           const int32_t* tile_offsets, const int32_t* flatten_ids,
                                           uint32_t n_isects, uint32_t C, uint32_t N,
                                           uint32_t width, uint32_t height, uint32_t tile_size,
                                           uint32_t n_tiles_x, uint32_t n_tiles_y, void* render,
                                           r2s_stream_t stream) {
    return launch_fwd(dtype, false, false, means2d, conics, opacities, mus, nullptr, nullptr,
                      tile_offsets, flatten_ids, n_isects, C, N, width, height, tile_size,
                      n_tiles_x, n_tiles_y, render, nullptr, stream);
}

extern "C" int r2s_rasterize_to_pixels_mean_var_fwd(
    int dtype, const void* means2d, const void* conics, const void* opacities, const void* mus,
    const void* uncertainties, const int32_t* tile_offsets, const int32_t* flatten_ids,
    uint32_t n_isects, uint32_t C, uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size,
    uint32_t n_tiles_x, uint32_t n_tiles_y, void* render, void* render_var, r2s_stream_t stream) {
    return launch_fwd(dtype, false, true, means2d, conics, opacities, mus, nullptr, uncertainties,
                      tile_offsets, flatten_ids, n_isects, C, N, width, height, tile_size,
                      n_tiles_x, n_tiles_y, render, render_var, stream);
}

extern "C" int r2s_beta_rasterize_to_pixels_fwd(int dtype, const void* means2d, const void* conics,
                                                const void* opacities, const void* mus,
                                                const void* betas, const int32_t* tile_offsets,
                                                const int32_t* flatten_ids, uint32_t n_isects,
                                                uint32_t C, uint32_t N, uint32_t width,
                                                uint32_t height, uint32_t tile_size,
                                                uint32_t n_tiles_x, uint32_t n_tiles_y,
                                                void* render, r2s_stream_t stream) {
    return launch_fwd(dtype, true, false, means2d, conics, opacities, mus, betas, nullptr,
                      tile_offsets, flatten_ids, n_isects, C, N, width, height, tile_size,
                      n_tiles_x, n_tiles_y, render, nullptr, stream);
}

extern "C" int r2s_beta_rasterize_to_pixels_mean_var_fwd(
    int dtype, const void* means2d, const void* conics, const void* opacities, const void* mus,
    const void* betas, const void* uncertainties, const int32_t* tile_offsets,
    const int32_t* flatten_ids, uint32_t n_isects, uint32_t C, uint32_t N, uint32_t width,
    uint32_t height, uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y, void* render,
    void* render_var, r2s_stream_t stream) {
    return launch_fwd(dtype, true, true, means2d, conics, opacities, mus, betas, uncertainties,
                      tile_offsets, flatten_ids, n_isects, C, N, width, height, tile_size,
                      n_tiles_x, n_tiles_y, render, render_var, stream);
}

extern "C" int r2s_rasterize_to_pixels_uncertainty_bwd(
    int dtype, const void* means2d, const void* conics, const void* opacities, const void* mus,
    const void* betas, const int32_t* tile_offsets, const int32_t* flatten_ids, uint32_t n_isects,
    uint32_t C, uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size,
    uint32_t n_tiles_x, uint32_t n_tiles_y, const void* v_render_var, void* v_uncertainties,
    r2s_stream_t stream) {
    if (int rc = check_tile(tile_size, n_tiles_x, n_tiles_y, width, height)) return rc;
    const bool beta = betas != nullptr;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (int rc = fill_zero<S>((S*)v_uncertainties, N, stream)) return rc;
        const uint64_t blocks = uint64_t(C) * n_tiles_x * n_tiles_y;
        if (blocks == 0 || n_isects == 0) return R2S_OK;
        const uint32_t threads = block_threads(tile_size);
        const size_t shm = lds_bytes(dtype, std::min(threads, kBatch));
        auto A = make_args<S>(means2d, conics, opacities, mus, betas, nullptr, tile_offsets,
                              flatten_ids, n_isects, C, N, width, height, tile_size, n_tiles_x,
                              n_tiles_y);
        auto k = beta ? raster_bwd_tile_kernel<S, true, 1> : raster_bwd_tile_kernel<S, false, 1>;
        k<<<uint32_t(blocks), threads, shm, stream>>>(A, (const S*)v_render_var, nullptr, nullptr,
                                                      nullptr, nullptr, nullptr, nullptr,
                                                      (S*)v_uncertainties);
        return check_launch("rasterize_to_pixels_uncertainty_bwd");
    });
}

extern "C" int r2s_rasterize_to_pixels_bwd(
    int dtype, const void* means2d, const void* conics, const void* opacities, const void* mus,
    const void* betas, const int32_t* tile_offsets, const int32_t* flatten_ids, uint32_t n_isects,
    uint32_t C, uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size,
    uint32_t n_tiles_x, uint32_t n_tiles_y, const void* v_render, void* v_means2d_abs,
    void* v_means2d, void* v_conics, void* v_opacities, void* v_mus, void* v_betas,
    r2s_stream_t stream) {
    if (int rc = check_tile(tile_size, n_tiles_x, n_tiles_y, width, height)) return rc;
    const bool beta = betas != nullptr;
    R2S_REQUIRE(!beta || v_betas != nullptr, "v_betas is required with betas");
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        const size_t CN = size_t(C) * N;
        if (int rc = zero_grads_2d<S>(CN, v_means2d_abs, v_means2d, v_conics, v_opacities, v_mus,
                                      beta ? v_betas : nullptr, stream))
            return rc;
        const uint64_t blocks = uint64_t(C) * n_tiles_x * n_tiles_y;
        if (blocks == 0 || n_isects == 0 || CN == 0) return R2S_OK;
        const uint32_t threads = block_threads(tile_size);
        const size_t shm = lds_bytes(dtype, std::min(threads, kBatch));
        auto A = make_args<S>(means2d, conics, opacities, mus, betas, nullptr, tile_offsets,
                              flatten_ids, n_isects, C, N, width, height, tile_size, n_tiles_x,
                              n_tiles_y);
        auto k = beta ? raster_bwd_tile_kernel<S, true, 0> : raster_bwd_tile_kernel<S, false, 0>;
        k<<<uint32_t(blocks), threads, shm, stream>>>(A, (const S*)v_render, (S*)v_means2d_abs,
                                                      (S*)v_means2d, (S*)v_conics, (S*)v_opacities,
                                                      (S*)v_mus, (S*)v_betas, nullptr);
        return check_launch("rasterize_to_pixels_bwd");
    });
}

extern "C" int r2s_rasterize_to_pixels_bwd_per_gaussian(
    int dtype, const void* means2d, const void* conics, const void* opacities, const void* mus,
    const void* betas, const int64_t* isect_ids, const int32_t* flatten_ids, uint32_t n_isects,
    uint32_t C, uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size,
    uint32_t n_tiles_x, uint32_t n_tiles_y, const void* v_render, const void* v_render_var,
    const void* uncertainties, void* v_means2d_abs, void* v_means2d, void* v_conics,
    void* v_opacities, void* v_mus, void* v_betas, void* v_uncertainties, r2s_stream_t stream) {
    if (int rc = check_tile(tile_size, n_tiles_x, n_tiles_y, width, height)) return rc;
    const bool beta = betas != nullptr;
    R2S_REQUIRE(!beta || v_betas != nullptr, "v_betas is required with betas");
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        const size_t CN = size_t(C) * N;
        if (int rc = zero_grads_2d<S>(CN, v_means2d_abs, v_means2d, v_conics, v_opacities, v_mus,
                                      beta ? v_betas : nullptr, stream))
            return rc;
        if (v_uncertainties)
            if (int rc = fill_zero<S>((S*)v_uncertainties, N, stream)) return rc;
        if (n_isects == 0 || CN == 0) return R2S_OK;
        auto A = make_args<S>(means2d, conics, opacities, mus, betas, nullptr, nullptr, flatten_ids,
                              n_isects, C, N, width, height, tile_size, n_tiles_x, n_tiles_y);
        constexpr uint32_t waves = 4;
        auto k = beta ? raster_bwd_isect_kernel<S, true> : raster_bwd_isect_kernel<S, false>;
        k<<<std::min<uint32_t>(ceil_div(n_isects, waves), kIsectBlocks), waves * kWave, 0, stream>>>(
            A, isect_ids, (const S*)v_render, (const S*)v_render_var,
            v_render_var ? (const S*)uncertainties : nullptr, (S*)v_means2d_abs, (S*)v_means2d,
            (S*)v_conics, (S*)v_opacities, (S*)v_mus, (S*)v_betas,
            v_render_var ? (S*)v_uncertainties : nullptr);
        return check_launch("rasterize_to_pixels_bwd_per_gaussian");
    });
}

// 3D voxelization: each voxel sums opacity * profile over the primitives binned
// to its 3D tile. Tile-parallel forward / VJP (LDS-staged batches) and the
// intersection-parallel VJP (one wave64 per intersection), for any tile size
// with tile_size^3 <= 1024 and both float/double.
// Profiles follow r2splat/cuda/csrc/VoxelizeToVoxelsFwd.cu:181-219 (Gaussian,
// threshold 1e-6) and the Beta kernels (threshold 1e-5); the VJPs follow
// VoxelizeToVoxelsBwd.cu:387-434 and the Beta per-primitive kernel.
#include "common.hpp"
#include "raster_common.hpp"

namespace r2s {

constexpr uint32_t kVoxBatch = 256;

template <typename S>
struct Prim3D {
    S x, y, z, c[6], opac, beta;
};

template <typename S, bool BETA>
__device__ __forceinline__ bool eval_profile3(const Prim3D<S>& p, S px, S py, S pz, S& vis, S& alpha) {
    if (!BETA) {
        const S dx = px - p.x, dy = py - p.y, dz = pz - p.z;
        const S sigma = S(0.5) * (p.c[0] * dx * dx + p.c[3] * dy * dy + p.c[5] * dz * dz) +
                        p.c[1] * dx * dy + p.c[2] * dx * dz + p.c[4] * dy * dz;
        vis = Math<S>::exp(-sigma);
        alpha = p.opac * vis;
        return !(sigma < S(0) || fabs(alpha) < S(1e-6));
    } else {
        const S dx = p.x - px, dy = p.y - py, dz = p.z - pz;
        const S q = p.c[0] * dx * dx + p.c[3] * dy * dy + p.c[5] * dz * dz +
                    S(2) * (p.c[1] * dx * dy + p.c[2] * dx * dz + p.c[4] * dy * dz);
        if (q < S(0) || q >= S(1)) return false;
        const S h = S(1) - q;
        vis = Math<S>::pow(h, p.beta);
        alpha = p.opac * vis;
        return !(fabs(alpha) < S(1e-5));
    }
}

template <typename S>
struct Grad3D {
    S xyz[3] = {S(0), S(0), S(0)};
    S c[6] = {S(0), S(0), S(0), S(0), S(0), S(0)};
    S opac = S(0), beta = S(0);

    __device__ __forceinline__ void reduce_wave() {
#pragma unroll
        for (int k = 0; k < 3; ++k) xyz[k] = wave_sum(xyz[k]);
#pragma unroll
        for (int k = 0; k < 6; ++k) c[k] = wave_sum(c[k]);
        opac = wave_sum(opac);
        beta = wave_sum(beta);
    }
    __device__ __forceinline__ void commit(int32_t g, S* v_xyz, S* v_c, S* v_opac, S* v_beta) const {
        const size_t G = size_t(g);
#pragma unroll
        for (int k = 0; k < 3; ++k) atomic_add(v_xyz + 3 * G + k, xyz[k]);
#pragma unroll
        for (int k = 0; k < 6; ++k) atomic_add(v_c + 6 * G + k, c[k]);
        atomic_add(v_opac + G, opac);
        if (v_beta) atomic_add(v_beta + G, beta);
    }
};

template <typename S, bool BETA>
__device__ __forceinline__ void voxel_vjp(const Prim3D<S>& p, S px, S py, S pz, S vis, S alpha, S v,
                                          Grad3D<S>& g) {
    if (!BETA) {
        const S dx = px - p.x, dy = py - p.y, dz = pz - p.z;
        const S vs = -v * alpha;  // d/d sigma
        g.c[0] += vs * S(0.5) * dx * dx;
        g.c[1] += vs * dx * dy;
        g.c[2] += vs * dx * dz;
        g.c[3] += vs * S(0.5) * dy * dy;
        g.c[4] += vs * dy * dz;
        g.c[5] += vs * S(0.5) * dz * dz;
        const S va = v * alpha;  // d sigma / d mean = -P delta
        g.xyz[0] += va * (p.c[0] * dx + p.c[1] * dy + p.c[2] * dz);
        g.xyz[1] += va * (p.c[1] * dx + p.c[3] * dy + p.c[4] * dz);
        g.xyz[2] += va * (p.c[2] * dx + p.c[4] * dy + p.c[5] * dz);
    } else {
        const S dx = p.x - px, dy = p.y - py, dz = p.z - pz;
        const S q = p.c[0] * dx * dx + p.c[3] * dy * dy + p.c[5] * dz * dz +
                    S(2) * (p.c[1] * dx * dy + p.c[2] * dx * dz + p.c[4] * dy * dz);
        const S h = S(1) - q;
        const S vq = -p.opac * p.beta * (vis / h) * v;  // d/d q
        g.c[0] += vq * dx * dx;
        g.c[1] += S(2) * vq * dx * dy;
        g.c[2] += S(2) * vq * dx * dz;
        g.c[3] += vq * dy * dy;
        g.c[4] += S(2) * vq * dy * dz;
        g.c[5] += vq * dz * dz;
        g.xyz[0] += S(2) * vq * (p.c[0] * dx + p.c[1] * dy + p.c[2] * dz);
        g.xyz[1] += S(2) * vq * (p.c[1] * dx + p.c[3] * dy + p.c[4] * dz);
        g.xyz[2] += S(2) * vq * (p.c[2] * dx + p.c[4] * dy + p.c[5] * dz);
        g.beta += v * alpha * Math<S>::log(h);
    }
    g.opac += v * vis;
}

template <typename S>
struct VoxArgs {
    const S* means3d; const S* conics; const S* opac; const S* betas; const S* unc;
    const int32_t* offsets; const int32_t* flatten; uint32_t n_isects, N, dx, dy, dz, T, ntx, nty, ntz;
};

template <typename S>
struct VoxLDS {
    S* f;  // [11][batch]: x y z c0..c5 opac beta
    S* q;
    int32_t* id;
    uint32_t batch;
    __device__ __forceinline__ Prim3D<S> prim(uint32_t t, bool beta) const {
        Prim3D<S> p;
        p.x = f[t]; p.y = f[batch + t]; p.z = f[2 * batch + t];
#pragma unroll
        for (int k = 0; k < 6; ++k) p.c[k] = f[(3 + k) * batch + t];
        p.opac = f[9 * batch + t];
        p.beta = beta ? f[10 * batch + t] : S(0);
        return p;
    }
};

inline size_t vox_lds_bytes(int dtype, uint32_t batch) {
    const size_t s = dtype == R2S_F64 ? 8 : 4;
    return batch * (12 * s + 4);
}

template <typename S>
__device__ __forceinline__ VoxLDS<S> vox_carve(unsigned char* smem, uint32_t batch) {
    VoxLDS<S> L;
    L.f = reinterpret_cast<S*>(smem);
    L.q = L.f + 11 * batch;
    L.id = reinterpret_cast<int32_t*>(L.f + 12 * batch);
    L.batch = batch;
    return L;
}

template <typename S, bool BETA, bool VAR>
__device__ __forceinline__ void vox_stage(const VoxArgs<S>& A, const VoxLDS<S>& L, int32_t b0, int32_t end) {
    const uint32_t t = threadIdx.x;
    if (t < L.batch && b0 + int32_t(t) < end) {
        int32_t g = A.flatten[b0 + t];
        const bool ok = g >= 0 && uint32_t(g) < A.N;
        if (!ok) g = 0;  // out-of-range ids contribute nothing
        const size_t G = size_t(g);
        L.id[t] = g;
        L.f[t] = A.means3d[3 * G];
        L.f[L.batch + t] = A.means3d[3 * G + 1];
        L.f[2 * L.batch + t] = A.means3d[3 * G + 2];
#pragma unroll
        for (int k = 0; k < 6; ++k) L.f[(3 + k) * L.batch + t] = A.conics[6 * G + k];
        L.f[9 * L.batch + t] = ok ? A.opac[g] : S(0);
        if (BETA) L.f[10 * L.batch + t] = A.betas[g];
        if (VAR) L.q[t] = A.unc[g];
    }
}

template <typename S>
__device__ __forceinline__ void vox_tile(const VoxArgs<S>& A, uint32_t blk, uint32_t& i, uint32_t& j,
                                         uint32_t& k, bool& inside, int32_t& start, int32_t& end) {
    const uint32_t tx = blk % A.ntx, rest = blk / A.ntx;
    const uint32_t ty = rest % A.nty, tz = rest / A.nty;
    const uint32_t t = threadIdx.x, T = A.T;
    const uint32_t lx = t % T, ly = (t / T) % T, lz = t / (T * T);
    k = tx * T + lx;
    j = ty * T + ly;
    i = tz * T + lz;
    inside = t < T * T * T && i < A.dz && j < A.dy && k < A.dx;
    start = A.offsets[blk];
    const uint32_t nt = A.ntx * A.nty * A.ntz;
    end = (blk + 1 == nt) ? int32_t(A.n_isects) : A.offsets[blk + 1];
    start = max(0, min(start, int32_t(A.n_isects)));
    end = max(start, min(end, int32_t(A.n_isects)));
}

template <typename S, bool BETA, bool VAR>
__global__ void vox_fwd_kernel(VoxArgs<S> A, S* __restrict__ out, S* __restrict__ out_var) {
    extern __shared__ __align__(16) unsigned char smem[];
    VoxLDS<S> L = vox_carve<S>(smem, min(blockDim.x, kVoxBatch));
    uint32_t i, j, k;
    bool inside;
    int32_t start, end;
    vox_tile(A, blockIdx.x, i, j, k, inside, start, end);
    const S px = S(k) + S(0.5), py = S(j) + S(0.5), pz = S(i) + S(0.5);
    S acc = S(0), acc_var = S(0);
    for (int32_t b0 = start; b0 < end; b0 += int32_t(L.batch)) {
        __syncthreads();
        vox_stage<S, BETA, VAR>(A, L, b0, end);
        __syncthreads();
        if (!inside) continue;
        const uint32_t cnt = uint32_t(min<int32_t>(int32_t(L.batch), end - b0));
        for (uint32_t t = 0; t < cnt; ++t) {
            const Prim3D<S> p = L.prim(t, BETA);
            S vis, alpha;
            if (!eval_profile3<S, BETA>(p, px, py, pz, vis, alpha)) continue;
            acc += alpha;
            if (VAR) acc_var += L.q[t] * alpha * alpha;
        }
    }
    if (inside) {
        const size_t o = (size_t(i) * A.dy + j) * A.dx + k;
        out[o] = acc;
        if (VAR) out_var[o] = acc_var;
    }
}

template <typename S, bool BETA>
__global__ void vox_bwd_tile_kernel(VoxArgs<S> A, const S* __restrict__ v_render, S* __restrict__ v_xyz,
                                    S* __restrict__ v_c, S* __restrict__ v_opac, S* __restrict__ v_beta) {
    extern __shared__ __align__(16) unsigned char smem[];
    VoxLDS<S> L = vox_carve<S>(smem, min(blockDim.x, kVoxBatch));
    uint32_t i, j, k;
    bool inside;
    int32_t start, end;
    vox_tile(A, blockIdx.x, i, j, k, inside, start, end);
    const S px = S(k) + S(0.5), py = S(j) + S(0.5), pz = S(i) + S(0.5);
    const S v = inside ? v_render[(size_t(i) * A.dy + j) * A.dx + k] : S(0);
    for (int32_t b0 = start; b0 < end; b0 += int32_t(L.batch)) {
        __syncthreads();
        vox_stage<S, BETA, false>(A, L, b0, end);
        __syncthreads();
        const uint32_t cnt = uint32_t(min<int32_t>(int32_t(L.batch), end - b0));
        for (uint32_t t = 0; t < cnt; ++t) {
            const Prim3D<S> p = L.prim(t, BETA);
            S vis = S(0), alpha = S(0);
            Grad3D<S> g;
            if (inside && eval_profile3<S, BETA>(p, px, py, pz, vis, alpha))
                voxel_vjp<S, BETA>(p, px, py, pz, vis, alpha, v, g);
            g.reduce_wave();
            if (lane_id() == 0) g.commit(L.id[t], v_xyz, v_c, v_opac, BETA ? v_beta : nullptr);
        }
    }
}

template <typename S, bool BETA>
__global__ void __launch_bounds__(256)
vox_bwd_isect_kernel(VoxArgs<S> A, const int64_t* __restrict__ isect_ids, const S* __restrict__ v_render,
                     const S* __restrict__ v_var, const S* __restrict__ unc, S* __restrict__ v_xyz,
                     S* __restrict__ v_c, S* __restrict__ v_opac, S* __restrict__ v_beta,
                     S* __restrict__ v_q) {
    // grid-stride over intersections (one wave each), at most kIsectBlocks blocks
    const uint32_t waves = blockDim.x / kWave;
    const uint64_t stride = uint64_t(gridDim.x) * waves;
    for (uint64_t e = uint64_t(blockIdx.x) * waves + threadIdx.x / kWave; e < A.n_isects; e += stride) {
        const uint32_t nt = A.ntx * A.nty * A.ntz;
        const int64_t key = isect_ids[e];
        if (key < 0 || key >= int64_t(nt)) continue;
        const uint32_t tile = uint32_t(key);
        const uint32_t tx = tile % A.ntx, ty = (tile / A.ntx) % A.nty, tz = tile / (A.ntx * A.nty);
        const int32_t g = A.flatten[e];
        if (g < 0 || uint32_t(g) >= A.N) continue;
        const size_t G = size_t(g);
        Prim3D<S> p;
        p.x = A.means3d[3 * G];
        p.y = A.means3d[3 * G + 1];
        p.z = A.means3d[3 * G + 2];
#pragma unroll
        for (int k = 0; k < 6; ++k) p.c[k] = A.conics[6 * G + k];
        p.opac = A.opac[g];
        p.beta = BETA ? A.betas[g] : S(0);
        const S q = unc ? unc[g] : S(0);
        Grad3D<S> gr;
        S vq = S(0);
        const uint32_t T = A.T, nvox = T * T * T;
        for (uint32_t t = lane_id(); t < nvox; t += kWave) {
            const uint32_t lx = t % T, ly = (t / T) % T, lz = t / (T * T);
            const uint32_t k = tx * T + lx, j = ty * T + ly, i = tz * T + lz;
            if (i >= A.dz || j >= A.dy || k >= A.dx) continue;
            const S px = S(k) + S(0.5), py = S(j) + S(0.5), pz = S(i) + S(0.5);
            S vis, alpha;
            if (!eval_profile3<S, BETA>(p, px, py, pz, vis, alpha)) continue;
            const size_t o = (size_t(i) * A.dy + j) * A.dx + k;
            S v = v_render[o];
            if (v_var) {
                const S vv = v_var[o];
                vq += vv * alpha * alpha;
                if (unc) v += S(2) * q * alpha * vv;
            }
            voxel_vjp<S, BETA>(p, px, py, pz, vis, alpha, v, gr);
        }
        gr.reduce_wave();
        if (v_q) vq = wave_sum(vq);
        if (lane_id() == 0) {
            gr.commit(g, v_xyz, v_c, v_opac, BETA ? v_beta : nullptr);
            if (v_q) atomic_add(v_q + G, vq);
        }
    }
}

inline int check_vox_tile(uint32_t T, uint32_t ntx, uint32_t nty, uint32_t ntz, uint32_t dx,
                          uint32_t dy, uint32_t dz) {
    R2S_REQUIRE(T >= 1 && T * T * T <= 1024, "tile_size %u must satisfy 1 <= tile_size^3 <= 1024", T);
    R2S_REQUIRE(uint64_t(ntx) * T >= dx && uint64_t(nty) * T >= dy && uint64_t(ntz) * T >= dz,
                "tile grid does not cover the volume");
    return R2S_OK;
}

inline uint32_t vox_threads(uint32_t T) { return ((T * T * T + kWave - 1) / kWave) * kWave; }

template <typename S>
VoxArgs<S> make_vox_args(const void* means3d, const void* conics, const void* opac, const void* betas,
                         const void* unc, const int32_t* offsets, const int32_t* flatten,
                         uint32_t n_isects, uint32_t N, uint32_t dx, uint32_t dy, uint32_t dz,
                         uint32_t T, uint32_t ntx, uint32_t nty, uint32_t ntz) {
    return VoxArgs<S>{(const S*)means3d, (const S*)conics, (const S*)opac, (const S*)betas,
                      (const S*)unc, offsets, flatten, n_isects, N, dx, dy, dz, T, ntx, nty, ntz};
}

}  // namespace r2s

using namespace r2s;

extern "C" int r2s_voxelize_to_voxels_fwd(int dtype, const void* means3d, const void* conics,
                                          const void* opacities, const void* betas,
                                          const void* uncertainties, const int32_t* tile_offsets,
                                          const int32_t* flatten_ids, uint32_t n_isects, uint32_t N,
                                          uint32_t dim_x, uint32_t dim_y, uint32_t dim_z,
                                          uint32_t tile_size, uint32_t n_tiles_x,
                                          uint32_t n_tiles_y, uint32_t n_tiles_z, void* render,
                                          void* render_var, r2s_stream_t stream) {
    if (int rc = check_vox_tile(tile_size, n_tiles_x, n_tiles_y, n_tiles_z, dim_x, dim_y, dim_z))
        return rc;
    const uint64_t blocks = uint64_t(n_tiles_x) * n_tiles_y * n_tiles_z;
    if (blocks == 0) return R2S_OK;
    const bool beta = betas != nullptr, var = render_var != nullptr;
    R2S_REQUIRE(!var || uncertainties != nullptr, "render_var requires uncertainties");
    const uint32_t threads = vox_threads(tile_size);
    const size_t shm = vox_lds_bytes(dtype, std::min(threads, kVoxBatch));
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        auto A = make_vox_args<S>(means3d, conics, opacities, betas, uncertainties, tile_offsets,
                                  flatten_ids, n_isects, N, dim_x, dim_y, dim_z, tile_size,
                                  n_tiles_x, n_tiles_y, n_tiles_z);
        auto k = beta ? (var ? vox_fwd_kernel<S, true, true> : vox_fwd_kernel<S, true, false>)
                      : (var ? vox_fwd_kernel<S, false, true> : vox_fwd_kernel<S, false, false>);
        k<<<uint32_t(blocks), threads, shm, stream>>>(A, (S*)render, (S*)render_var);
        return check_launch("voxelize_to_voxels_fwd");
    });
}

extern "C" int r2s_voxelize_to_voxels_bwd(int dtype, const void* means3d, const void* conics,
                                          const void* opacities, const void* betas,
                                          const int32_t* tile_offsets, const int32_t* flatten_ids,
                                          uint32_t n_isects, uint32_t N, uint32_t dim_x,
                                          uint32_t dim_y, uint32_t dim_z, uint32_t tile_size,
                                          uint32_t n_tiles_x, uint32_t n_tiles_y,
                                          uint32_t n_tiles_z, const void* v_render, void* v_means3d,
                                          void* v_conics, void* v_opacities, void* v_betas,
                                          r2s_stream_t stream) {
    if (int rc = check_vox_tile(tile_size, n_tiles_x, n_tiles_y, n_tiles_z, dim_x, dim_y, dim_z))
        return rc;
    const bool beta = betas != nullptr;
    R2S_REQUIRE(!beta || v_betas != nullptr, "v_betas is required with betas");
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (int rc = fill_zero<S>((S*)v_means3d, 3 * size_t(N), stream)) return rc;
        if (int rc = fill_zero<S>((S*)v_conics, 6 * size_t(N), stream)) return rc;
        if (int rc = fill_zero<S>((S*)v_opacities, N, stream)) return rc;
        if (beta)
            if (int rc = fill_zero<S>((S*)v_betas, N, stream)) return rc;
        const uint64_t blocks = uint64_t(n_tiles_x) * n_tiles_y * n_tiles_z;
        if (blocks == 0 || n_isects == 0 || N == 0) return R2S_OK;
        const uint32_t threads = vox_threads(tile_size);
        const size_t shm = vox_lds_bytes(dtype, std::min(threads, kVoxBatch));
        auto A = make_vox_args<S>(means3d, conics, opacities, betas, nullptr, tile_offsets,
                                  flatten_ids, n_isects, N, dim_x, dim_y, dim_z, tile_size,
                                  n_tiles_x, n_tiles_y, n_tiles_z);
        auto k = beta ? vox_bwd_tile_kernel<S, true> : vox_bwd_tile_kernel<S, false>;
        k<<<uint32_t(blocks), threads, shm, stream>>>(A, (const S*)v_render, (S*)v_means3d,
                                                      (S*)v_conics, (S*)v_opacities, (S*)v_betas);
        return check_launch("voxelize_to_voxels_bwd");
    });
}

extern "C" int r2s_voxelize_to_voxels_bwd_per_gaussian(
    int dtype, const void* means3d, const void* conics, const void* opacities, const void* betas,
    const int64_t* isect_ids, const int32_t* flatten_ids, uint32_t n_isects, uint32_t N,
    uint32_t dim_x, uint32_t dim_y, uint32_t dim_z, uint32_t tile_size, uint32_t n_tiles_x,
    uint32_t n_tiles_y, uint32_t n_tiles_z, const void* v_render, const void* v_render_var,
    const void* uncertainties, void* v_means3d, void* v_conics, void* v_opacities, void* v_betas,
    void* v_uncertainties, r2s_stream_t stream) {
    if (int rc = check_vox_tile(tile_size, n_tiles_x, n_tiles_y, n_tiles_z, dim_x, dim_y, dim_z))
        return rc;
    const bool beta = betas != nullptr;
    R2S_REQUIRE(!beta || v_betas != nullptr, "v_betas is required with betas");
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (int rc = fill_zero<S>((S*)v_means3d, 3 * size_t(N), stream)) return rc;
        if (int rc = fill_zero<S>((S*)v_conics, 6 * size_t(N), stream)) return rc;
        if (int rc = fill_zero<S>((S*)v_opacities, N, stream)) return rc;
        if (beta)
            if (int rc = fill_zero<S>((S*)v_betas, N, stream)) return rc;
        if (v_uncertainties)
            if (int rc = fill_zero<S>((S*)v_uncertainties, N, stream)) return rc;
        if (n_isects == 0 || N == 0) return R2S_OK;
        auto A = make_vox_args<S>(means3d, conics, opacities, betas, nullptr, nullptr, flatten_ids,
                                  n_isects, N, dim_x, dim_y, dim_z, tile_size, n_tiles_x,
                                  n_tiles_y, n_tiles_z);
        constexpr uint32_t waves = 4;
        auto k = beta ? vox_bwd_isect_kernel<S, true> : vox_bwd_isect_kernel<S, false>;
        k<<<std::min<uint32_t>(ceil_div(n_isects, waves), kIsectBlocks), waves * kWave, 0, stream>>>(
            A, isect_ids, (const S*)v_render, (const S*)v_render_var,
            v_render_var ? (const S*)uncertainties : nullptr, (S*)v_means3d, (S*)v_conics,
            (S*)v_opacities, (S*)v_betas, v_render_var ? (S*)v_uncertainties : nullptr);
        return check_launch("voxelize_to_voxels_bwd_per_gaussian");
    });
}

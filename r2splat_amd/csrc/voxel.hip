// 3D voxelization: each voxel sums opacity * profile over the primitives binned
// to its 3D tile. Tile-parallel forward / VJP (LDS-staged batches) and the
// intersection-parallel VJP (one wave64 per intersection), for any tile size
// with tile_size^3 <= 1024 and both float/double.
// Profiles follow r2splat/cuda/csrc/VoxelizeToVoxelsFwd.cu:181-219 (Gaussian,
// threshold 1e-6) and the Beta kernels (threshold 1e-5); the VJPs follow
// VoxelizeToVoxelsBwd.cu:387-434 and the Beta per-primitive kernel.
#include <type_traits>

#include "common.hpp"
#include "binning.hpp"
#include "geometry_common.hpp"
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
    // fused float32 kernels: each primitive's voxel-frame mean and inverse
    // covariance in double precision (kFrame doubles per primitive)
    const double* frame = nullptr;
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

// ---------------------------------------------------------------------------
// Moments form of the per-primitive VJP (float32, mean only, 8^3 tiles): each
// workgroup takes a fixed chunk of the tile-sorted intersections (equal work
// per workgroup, whatever the tile populations), one per lane, and visits the
// chunk's tiles in order with each tile's adjoint staged in LDS; a lane sums its
// primitive's moments over the 512 voxels row by row (every lane of the tile
// reads the same row: broadcast loads, and no cross-lane reduction):
//   Gaussian, delta = voxel - mean, G = v e^-sigma (terms that pass the cut-off):
//     M0 = sum G, M_k = sum G delta_k, M_kl = sum G delta_k delta_l,
//     v_mean = opac P M, v_conic = -opac (M_xx/2, M_xy, M_xz, M_yy/2, M_yz, M_zz/2),
//     v_opac = M0;
//   Beta, h = 1 - q, G = v h^(beta - 1): v_mean = 2 opac beta P M,
//     v_conic = -opac beta (M_xx, 2M_xy, 2M_xz, M_yy, 2M_yz, M_zz),
//     v_opac = sum v h^beta, v_beta = opac ln2 sum v h^beta log2 h;
// the same VJP as voxel_vjp summed over the tile. sigma is evaluated in log2
// units (P scaled by log2 e) and the cut-off "sigma >= 0 and |opac e^-sigma| >=
// 1e-6" (Beta: "0 <= q < 1 and |opac h^beta| >= 1e-5") is one unsigned compare
// of the bits of q against qlim. Gradients are accumulated with one atomic per
// field.
// ---------------------------------------------------------------------------
constexpr uint32_t kMomT = 8;

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The fused float32 kernels read each primitive's voxel-frame mean and
// inverse covariance from a double-precision frame, kFrame doubles per
// primitive (mean x y z, then the conic xx xy xz yy yz zz), computed from the
// world-frame parameters in double (vox_frame_kernel: the operator chain
// parameter transform -> volume mapping, in double). Gaussian offsets are
// formed as (v - hi) - lo with hi + lo the mean split into two floats, so the
// float32 rounding of a voxel coordinate (up to 1.5e-5 voxels at 256) never
// enters the profile. The Beta profile (1 - d^2)^beta is steep at its support
// edge when beta < 1 (there a change of 1e-7 in d^2 moves a term by up to
// ~3e-4), so Beta kernels form d^2 and 1 - d^2 in double from the frame.
constexpr int kFrame = 9;

__global__ void __launch_bounds__(256)
vox_frame_kernel(const float* __restrict__ means, const float* __restrict__ quats, const float* __restrict__ scales,
                 const double* __restrict__ ranges, const double* __restrict__ centers, uint32_t dx, uint32_t dy,
                 uint32_t dz, double eps, uint32_t N, double* __restrict__ frame) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double raw[4] = {double(quats[4 * size_t(n)]), double(quats[4 * size_t(n) + 1]),
                           double(quats[4 * size_t(n) + 2]), double(quats[4 * size_t(n) + 3])};
    double q[4], mag, norm, R[9], cov[6];
    unit_quat(raw, q, mag, norm);
    quat_to_rot(q, R);
    const double s[3] = {double(scales[3 * size_t(n)]), double(scales[3 * size_t(n) + 1]),
                         double(scales[3 * size_t(n) + 2])};
    covariance_from(R, s, cov);
    const double D[3] = {double(dx), double(dy), double(dz)};
    const double r[3] = {D[0] / ranges[0], D[1] / ranges[1], D[2] / ranges[2]};
    double* f = frame + kFrame * size_t(n);
#pragma unroll
    for (int k = 0; k < 3; ++k) f[k] = (double(means[3 * size_t(n) + k]) - centers[k]) * r[k] + D[k] * 0.5;
    const double sv[6] = {r[0] * r[0] * cov[0] + eps, r[0] * r[1] * cov[1], r[0] * r[2] * cov[2],
                          r[1] * r[1] * cov[3] + eps, r[1] * r[2] * cov[4], r[2] * r[2] * cov[5] + eps};
    double inv[6];
    spd3_inverse(sv, inv);  // zeros when singular (such primitives have zero radii and no tiles)
#pragma unroll
    for (int k = 0; k < 6; ++k) f[3 + k] = inv[k];
}

// Mean of primitive G split into float hi + lo (lo = 0 without a frame).
__device__ __forceinline__ void vox_mean_split(const VoxArgs<float>& A, size_t G, float hi[3], float lo[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (A.frame) {
            const double m = A.frame[kFrame * G + k];
            hi[k] = float(m);
            lo[k] = float(m - double(hi[k]));
        } else {
            hi[k] = A.means3d[3 * G + k];
            lo[k] = 0.f;
        }
    }
}

// Conic of primitive G (from the frame when there is one).
__device__ __forceinline__ void vox_conic(const VoxArgs<float>& A, size_t G, float P[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) P[k] = A.frame ? float(A.frame[kFrame * G + 3 + k]) : A.conics[6 * G + k];
}

// Largest Beta d^2 whose term passes the cut-off, in double: |alpha| =
// |w| (1 - d^2)^beta >= 1e-5 for 0 <= d^2 <= qmax; negative when none does.
__device__ __forceinline__ double beta_qmax(float w, float beta) {
    const double aw = fabs(double(w));
    if (!(aw >= 1e-5) || !(beta > 0.f)) return -1.0;
    return 1.0 - pow(1e-5 / aw, 1.0 / double(beta));
}

constexpr uint32_t kVoxPartial = 12;  // floats per intersection partial (11, or 12 with a variance)
constexpr uint32_t kMomThreads = 256;
constexpr float kL2E = 1.4426950408889634f;

__device__ __forceinline__ uint32_t vox_qlim(float opac, float beta, bool is_beta) {
    if (!is_beta) {
        const float S = __log2f(fabsf(opac)) + 19.931568569324174f;  // log2(|opac| / 1e-6)
        return (opac != 0.f && S >= 0.f) ? __float_as_uint(S) + 1u : 0u;
    }
    const float S = __log2f(fabsf(opac)) + 16.609640474436812f;      // log2(|opac| / 1e-5)
    const float D = 1.f - __builtin_amdgcn_exp2f(-S / beta);
    if (!(opac != 0.f && S >= 0.f && D >= 0.f)) return 0u;
    return __float_as_uint(fminf(D, 0.99999994f)) + 1u;
}

// VAR: the voxel adjoint becomes v + var_scale q alpha v_var (var_scale 2, or
// 0 when the variance is detached) and the partial's twelfth slot carries
// v_q = sum v_var alpha^2 (variance = sum q alpha^2). WIDE (Beta with a
// frame): q and h = 1 - q are formed in double from the frame, with the
// cut-off 0 <= q <= qmax decided in double, exactly as the fused forward does.
template <bool BETA, bool VAR, bool WIDE = false>
__global__ void __launch_bounds__(kMomThreads)
vox_bwd_moments_kernel(VoxArgs<float> A, const int64_t* __restrict__ isect_ids, const float* __restrict__ v_render,
                       const float* __restrict__ v_render_var, float var_scale,
                       float* __restrict__ v_xyz, float* __restrict__ v_c, float* __restrict__ v_opac,
                       float* __restrict__ v_beta, const int32_t* __restrict__ radii,
                       const uint32_t* __restrict__ cum, float* __restrict__ partials) {
    constexpr uint32_t T = kMomT, T3 = T * T * T;
    __shared__ __align__(16) float sv[T3];
    __shared__ __align__(16) float svv[VAR ? T3 : 4];
    __shared__ int64_t ids[kMomThreads];
    // a chunk of kMomThreads consecutive (tile-sorted) intersections, one per thread; the chunk's
    // tiles are visited in order, each one's adjoint staged in LDS in turn
    const uint32_t e0 = blockIdx.x * kMomThreads, e1 = min(e0 + kMomThreads, A.n_isects);
    const uint32_t e = e0 + threadIdx.x;
    const int64_t mine = e < e1 ? isect_ids[e] : INT64_MAX;
    ids[threadIdx.x] = mine;
    __syncthreads();
    const uint32_t n_here = e1 - e0;
    const int64_t nt = int64_t(A.ntx) * A.nty * A.ntz;
    static_assert(!WIDE || BETA, "the double-precision form is for Beta profiles");
    uint32_t pos = 0;  // first chunk entry of the current tile
    while (pos < n_here) {
        const int64_t tile = ids[pos];
        uint32_t lo = pos, hi = n_here;  // first entry of the next tile
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ids[mid] <= tile) lo = mid + 1; else hi = mid;
        }
        const bool valid_tile = tile >= 0 && tile < nt;
        const uint32_t blk = valid_tile ? uint32_t(tile) : 0u;
        const uint32_t tx = blk % A.ntx, rest = blk / A.ntx, ty = rest % A.nty, tz = rest / A.nty;
        if (valid_tile)
            for (uint32_t p = threadIdx.x; p < T3; p += blockDim.x) {
                const uint32_t k = tx * T + p % T, j = ty * T + (p / T) % T, i = tz * T + p / (T * T);
                const bool in = i < A.dz && j < A.dy && k < A.dx;
                const size_t o = (size_t(i) * A.dy + j) * A.dx + k;
                sv[p] = in ? v_render[o] : 0.f;
                if constexpr (VAR) svv[p] = in ? v_render_var[o] : 0.f;
            }
        __syncthreads();
        if (valid_tile && mine == tile) {
            const float x0 = float(tx * T) + 0.5f, y0 = float(ty * T) + 0.5f, z0 = float(tz * T) + 0.5f;
            const int32_t g = A.flatten[e];
            if (g >= 0 && uint32_t(g) < A.N) {
                const size_t G = size_t(g);
                float hi3[3], lo3[3], P[6];
                vox_mean_split(A, G, hi3, lo3);
                vox_conic(A, G, P);
                const float mx = hi3[0], my = hi3[1], mz = hi3[2];
                const float opac = A.opac[g], beta = BETA ? A.betas[g] : 0.f;
                uint32_t qlim = vox_qlim(opac, beta, BETA);
                double md[3] = {0.0, 0.0, 0.0}, Pd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, qmax = -1.0;
                if constexpr (WIDE) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) md[k] = A.frame[kFrame * G + k];
#pragma unroll
                    for (int k = 0; k < 6; ++k) Pd[k] = A.frame[kFrame * G + 3 + k];
                    qmax = beta_qmax(opac, beta);
                    qlim = qmax >= 0.0 ? 1u : 0u;  // liveness only: the cut-off is q <= qmax
                }
                const double ad = Pd[0], bxyd = 2.0 * Pd[1], bxzd = 2.0 * Pd[2], cyyd = Pd[3], cyzd = 2.0 * Pd[4],
                             czzd = Pd[5];
                const double dx0d = double(float(tx * T) + 0.5f) - md[0];
                const float kvar = VAR ? var_scale * A.unc[g] * opac : 0.f;  // d variance / d (alpha / w) factor
                // q = a dx^2 + bxy dx dy + bxz dx dz + cyy dy^2 + cyz dy dz + czz dz^2
                const float s = BETA ? 1.f : 0.5f * kL2E, t2 = BETA ? 2.f : kL2E;
                const float a = s * P[0], bxy = t2 * P[1], bxz = t2 * P[2], cyy = s * P[3], cyz = t2 * P[4], czz = s * P[5];
                const float bm1 = beta - 1.f;
                float m0 = 0.f, m_x = 0.f, m_y = 0.f, m_z = 0.f, mxx = 0.f, myy = 0.f, mzz = 0.f, mxy = 0.f, mxz = 0.f,
                      myz = 0.f, se = 0.f, sl = 0.f, vq = 0.f;
                const float dx0 = (x0 - mx) - lo3[0];
                if (qlim != 0u) {
                    for (uint32_t z = 0; z < T; ++z) {
                        const double dzd = double(z0 + float(z)) - md[2];
                        const float dz = WIDE ? float(dzd) : ((z0 + float(z)) - mz) - lo3[2];
                        for (uint32_t y = 0; y < T; ++y) {
                            const double dyd = double(y0 + float(y)) - md[1];
                            const float dy = WIDE ? float(dyd) : ((y0 + float(y)) - my) - lo3[1];
                            const float4* row = reinterpret_cast<const float4*>(sv + (z * T + y) * T);
                            const float4 va = row[0], vb = row[1];
                            const float v[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
                            float vv[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                            if constexpr (VAR) {
                                const float4* rowv = reinterpret_cast<const float4*>(svv + (z * T + y) * T);
                                const float4 wa = rowv[0], wb = rowv[1];
                                vv[0] = wa.x; vv[1] = wa.y; vv[2] = wa.z; vv[3] = wa.w;
                                vv[4] = wb.x; vv[5] = wb.y; vv[6] = wb.z; vv[7] = wb.w;
                            }
                            const float c1 = fmaf(bxy, dy, bxz * dz);
                            const float c0 = fmaf(dy, fmaf(cyy, dy, cyz * dz), czz * dz * dz);
                            const double c1d = fma(bxyd, dyd, bxzd * dzd);
                            const double c0d = fma(dyd, fma(cyyd, dyd, cyzd * dzd), czzd * dzd * dzd);
                            float r0 = 0.f, r1 = 0.f, r2 = 0.f, re = 0.f, rl = 0.f;
#pragma unroll
                            for (int x = 0; x < 8; ++x) {
                                float dx, q, h;
                                bool in;
                                if constexpr (WIDE) {
                                    const double dxd = dx0d + double(x);
                                    const double qd = fma(dxd, fma(ad, dxd, c1d), c0d);
                                    in = qd >= 0.0 && qd <= qmax;
                                    h = float(1.0 - qd);
                                    q = 0.f;
                                    dx = float(dxd);
                                } else {
                                    dx = dx0 + float(x);
                                    q = fmaf(dx, fmaf(a, dx, c1), c0);
                                    in = __float_as_uint(q) < qlim;
                                    h = 1.f - q;
                                }
                                float gv;
                                if constexpr (BETA) {
                                    const float lg = __log2f(h);
                                    const float hb = in ? __builtin_amdgcn_exp2f(bm1 * lg) : 0.f;  // h^(beta - 1)
                                    float ve;
                                    if constexpr (VAR) {
                                        const float he = hb * h;  // h^beta = alpha / w
                                        const float vj = fmaf(kvar * he, vv[x], v[x]);
                                        gv = vj * hb;
                                        ve = vj * he;
                                        vq = fmaf(vv[x] * he, he, vq);
                                    } else {
                                        gv = v[x] * hb;
                                        ve = gv * h;
                                    }
                                    re += ve;
                                    rl = fmaf(ve, in ? lg : 0.f, rl);
                                } else {
                                    const float e = in ? __builtin_amdgcn_exp2f(-q) : 0.f;
                                    if constexpr (VAR) {
                                        const float vj = fmaf(kvar * e, vv[x], v[x]);
                                        gv = vj * e;
                                        vq = fmaf(vv[x] * e, e, vq);
                                    } else {
                                        gv = v[x] * e;
                                    }
                                }
                                r0 += gv;
                                const float gx = gv * dx;
                                r1 += gx;
                                r2 = fmaf(gx, dx, r2);
                            }
                            m0 += r0;
                            m_x += r1;
                            mxx += r2;
                            m_y = fmaf(dy, r0, m_y);
                            m_z = fmaf(dz, r0, m_z);
                            myy = fmaf(dy * dy, r0, myy);
                            mzz = fmaf(dz * dz, r0, mzz);
                            myz = fmaf(dy * dz, r0, myz);
                            mxy = fmaf(dy, r1, mxy);
                            mxz = fmaf(dz, r1, mxz);
                            if constexpr (BETA) {
                                se += re;
                                sl += rl;
                            }
                        }
                    }
                }
                const float k = BETA ? 2.f * opac * beta : opac;
                const float gm[3] = {k * (P[0] * m_x + P[1] * m_y + P[2] * m_z), k * (P[1] * m_x + P[3] * m_y + P[4] * m_z),
                                     k * (P[2] * m_x + P[4] * m_y + P[5] * m_z)};
                float gc[6];
                if constexpr (BETA) {
                    const float kb = -opac * beta;
                    gc[0] = kb * mxx; gc[1] = 2.f * kb * mxy; gc[2] = 2.f * kb * mxz;
                    gc[3] = kb * myy; gc[4] = 2.f * kb * myz; gc[5] = kb * mzz;
                } else {
                    gc[0] = -0.5f * opac * mxx; gc[1] = -opac * mxy; gc[2] = -opac * mxz;
                    gc[3] = -0.5f * opac * myy; gc[4] = -opac * myz; gc[5] = -0.5f * opac * mzz;
                }
                const float go = BETA ? se : m0, gb = BETA ? opac * 0.6931471805599453f * sl : 0.f;
                if (partials) {
                    // the partial of this intersection at its emission index: the binning wrote
                    // the primitive's tiles from cum[n] on, in z, y, x order over its tile box
                    uint32_t lo[3], hi[3];
                    if (box_3d(A.means3d, radii, G, T, A.ntx, A.nty, A.ntz, lo, hi) && tx >= lo[0] && tx < hi[0] &&
                        ty >= lo[1] && ty < hi[1] && tz >= lo[2] && tz < hi[2]) {
                        const uint32_t w = cum[G] + ((tz - lo[2]) * (hi[1] - lo[1]) + (ty - lo[1])) * (hi[0] - lo[0]) +
                                           (tx - lo[0]);
                        if (w < A.n_isects) {
                            float4* out = reinterpret_cast<float4*>(partials + kVoxPartial * size_t(w));
                            out[0] = make_float4(gm[0], gm[1], gm[2], gc[0]);
                            out[1] = make_float4(gc[1], gc[2], gc[3], gc[4]);
                            out[2] = make_float4(gc[5], go, gb, VAR ? opac * opac * vq : 0.f);
                        }
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < 3; ++q) atomic_add(v_xyz + 3 * G + q, gm[q]);
#pragma unroll
                    for (int q = 0; q < 6; ++q) atomic_add(v_c + 6 * G + q, gc[q]);
                    atomic_add(v_opac + G, go);
                    if constexpr (BETA) atomic_add(v_beta + G, gb);
                }
            }
        }
        __syncthreads();  // every thread is done with this tile's adjoint
        pos = lo;
    }
}

// ---------------------------------------------------------------------------
// Fused float32 forward (mean only, 8^3 tiles): kVoxFwdWaves waves per tile,
// wave k summing the k-th contiguous slice of the tile's list; lane = one
// (y, z) row of 8 voxels. Records (mean, the quadratic form in log2 units for
// Gaussians, qlim, opacity, beta) are built from the primitive arrays a batch
// of 64 ahead by the lanes and read back from LDS by broadcast; per record a
// lane forms the row terms once and evaluates q = a dx^2 + (bxy dy + bxz dz) dx
// + c(dy, dz) for its 8 voxels. The slices' partial volumes are summed in LDS
// in wave order, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int kVoxFwdWaves = 4;

// Gaussian record: the quadratic form in log2 units, the mean as hi + lo floats.
struct __align__(16) VoxRec {
    float mx, my, mz, a, bxy, bxz, cyy, cyz, czz, w, beta;
    uint32_t qlim;
    float lx, ly, lz, pad;  // the mean's float32 remainder (vox_mean_split)
};
static_assert(sizeof(VoxRec) == 64, "voxel record must be 64 bytes");

// Beta record: the mean and the conic (off-diagonals doubled) in double; a
// term passes the cut-off when 0 <= d^2 <= qmax (qlim = 1 when some term can).
struct __align__(16) VoxRecB {
    double mx, my, mz, a, bxy, bxz, cyy, cyz, czz, qmax;
    float w, beta;
    uint32_t qlim, pad;
};
static_assert(sizeof(VoxRecB) == 96, "Beta voxel record must be 96 bytes");

template <bool BETA>
using VoxRecOf = typename std::conditional<BETA, VoxRecB, VoxRec>::type;

__device__ __forceinline__ VoxRec load_gauss_rec(const VoxArgs<float>& A, int32_t i, int32_t hi) {
    VoxRec r = {};
    if (i >= hi) return r;
    const int32_t g = A.flatten[i];
    if (g < 0 || uint32_t(g) >= A.N) return r;
    const size_t G = size_t(g);
    float h3[3], l3[3], P[6];
    vox_mean_split(A, G, h3, l3);
    vox_conic(A, G, P);
    constexpr float s = 0.5f * kL2E, t2 = kL2E;
    r.mx = h3[0];
    r.my = h3[1];
    r.mz = h3[2];
    r.lx = l3[0];
    r.ly = l3[1];
    r.lz = l3[2];
    r.a = s * P[0];
    r.bxy = t2 * P[1];
    r.bxz = t2 * P[2];
    r.cyy = s * P[3];
    r.cyz = t2 * P[4];
    r.czz = s * P[5];
    r.w = A.opac[g];
    r.qlim = vox_qlim(r.w, 0.f, false);
    return r;
}

// Needs the frame (the fused entry points require it).
__device__ __forceinline__ VoxRecB load_beta_rec(const VoxArgs<float>& A, int32_t i, int32_t hi) {
    VoxRecB r = {};
    if (i >= hi) return r;
    const int32_t g = A.flatten[i];
    if (g < 0 || uint32_t(g) >= A.N) return r;
    const double* f = A.frame + kFrame * size_t(g);
    r.mx = f[0];
    r.my = f[1];
    r.mz = f[2];
    r.a = f[3];
    r.bxy = 2.0 * f[4];
    r.bxz = 2.0 * f[5];
    r.cyy = f[6];
    r.cyz = 2.0 * f[7];
    r.czz = f[8];
    r.w = A.opac[g];
    r.beta = A.betas[g];
    r.qmax = beta_qmax(r.w, r.beta);
    r.qlim = r.qmax >= 0.0 ? 1u : 0u;
    return r;
}

template <bool BETA>
__device__ __forceinline__ VoxRecOf<BETA> load_vox_rec(const VoxArgs<float>& A, int32_t i, int32_t hi) {
    if constexpr (BETA)
        return load_beta_rec(A, i, hi);
    else
        return load_gauss_rec(A, i, hi);
}

// The uncertainty of the primitive a list entry names (0 past the list).
__device__ __forceinline__ float load_vox_unc(const VoxArgs<float>& A, int32_t i, int32_t hi) {
    if (i >= hi) return 0.f;
    const int32_t g = A.flatten[i];
    return (g < 0 || uint32_t(g) >= A.N) ? 0.f : A.unc[g];
}

// VAR also accumulates the variance sum q alpha^2 (q the primitive's
// uncertainty, staged beside the record) and sums it in a second LDS phase.
// Beta terms form d^2 and 1 - d^2 in double (the record's mean and conic),
// decide the cut-off on d^2 in double and take the power in float32.
template <bool BETA, bool VAR>
__global__ void __launch_bounds__(kVoxFwdWaves * kWave)
vox_fwd_fused_kernel(VoxArgs<float> A, float* __restrict__ out, float* __restrict__ out_var) {
    constexpr uint32_t T = kMomT;
    using R = VoxRecOf<BETA>;
    __shared__ R stage[kVoxFwdWaves][kWave];
    __shared__ float stage_q[VAR ? kVoxFwdWaves : 1][kWave];
    __shared__ float part[kVoxFwdWaves][T * T * T];
    const uint32_t blk = blockIdx.x;
    const uint32_t nt = A.ntx * A.nty * A.ntz;
    int32_t start = A.offsets[blk];
    int32_t end = (blk + 1 == nt) ? int32_t(A.n_isects) : A.offsets[blk + 1];
    start = max(0, min(start, int32_t(A.n_isects)));
    end = max(start, min(end, int32_t(A.n_isects)));
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
    const int32_t per = (end - start + kVoxFwdWaves - 1) / kVoxFwdWaves;
    const int32_t lo = min(end, start + int32_t(wave) * per), hi = min(end, lo + per);
    const uint32_t tx = blk % A.ntx, rest = blk / A.ntx, ty = rest % A.nty, tz = rest / A.nty;
    const uint32_t ly = lane % T, lz = lane / T;
    const float py = float(ty * T + ly) + 0.5f, pz = float(tz * T + lz) + 0.5f, px0 = float(tx * T) + 0.5f;
    float acc[T], acc_var[T];
#pragma unroll
    for (uint32_t x = 0; x < T; ++x) {
        acc[x] = 0.f;
        acc_var[x] = 0.f;
    }
    R* slot = stage[wave];
    R next = load_vox_rec<BETA>(A, lo + int32_t(lane), hi);
    float next_q = VAR ? load_vox_unc(A, lo + int32_t(lane), hi) : 0.f;
    for (int32_t base = lo; base < hi; base += kWave) {
        wave_sync_lds();
        slot[lane] = next;
        if constexpr (VAR) stage_q[wave][lane] = next_q;
        wave_sync_lds();
        next = load_vox_rec<BETA>(A, base + kWave + int32_t(lane), hi);
        if constexpr (VAR) next_q = load_vox_unc(A, base + kWave + int32_t(lane), hi);
        const int32_t n = min(int32_t(kWave), hi - base);
        for (int32_t j = 0; j < n; ++j) {
            const R r = slot[j];
            if (r.qlim == 0u) continue;  // wave-uniform (every lane reads the same record)
            const float unc = VAR ? stage_q[VAR ? wave : 0][j] : 0.f;
            float ek[T];
            if constexpr (BETA) {
                const double dy = double(py) - r.my, dz = double(pz) - r.mz, dx0 = double(px0) - r.mx;
                const double c1 = fma(r.bxy, dy, r.bxz * dz);
                const double c0 = fma(dy, fma(r.cyy, dy, r.cyz * dz), r.czz * dz * dz);
#pragma unroll
                for (uint32_t x = 0; x < T; ++x) {
                    const double dx = dx0 + double(x);
                    const double q = fma(dx, fma(r.a, dx, c1), c0);
                    const float e = __builtin_amdgcn_exp2f(r.beta * __log2f(float(1.0 - q)));
                    ek[x] = (q >= 0.0 && q <= r.qmax) ? e : 0.f;
                }
            } else {
                const float dy = (py - r.my) - r.ly, dz = (pz - r.mz) - r.lz, dx0 = (px0 - r.mx) - r.lx;
                const float c1 = fmaf(r.bxy, dy, r.bxz * dz);
                const float c0 = fmaf(dy, fmaf(r.cyy, dy, r.cyz * dz), r.czz * dz * dz);
#pragma unroll
                for (uint32_t x = 0; x < T; ++x) {
                    const float dx = dx0 + float(x);
                    const float q = fmaf(dx, fmaf(r.a, dx, c1), c0);
                    const float e = __builtin_amdgcn_exp2f(-q);
                    ek[x] = __float_as_uint(q) < r.qlim ? e : 0.f;
                }
            }
#pragma unroll
            for (uint32_t x = 0; x < T; ++x) {
                if constexpr (VAR) {
                    const float alpha = ek[x] * r.w;
                    acc[x] += alpha;
                    acc_var[x] = fmaf(unc * alpha, alpha, acc_var[x]);
                } else {
                    acc[x] = fmaf(ek[x], r.w, acc[x]);
                }
            }
        }
    }
#pragma unroll
    for (uint32_t x = 0; x < T; ++x) part[wave][(lz * T + ly) * T + x] = acc[x];
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < T * T * T; p += blockDim.x) {
        float v = 0.f;
#pragma unroll
        for (int k = 0; k < kVoxFwdWaves; ++k) v += part[k][p];
        const uint32_t k = tx * T + p % T, j = ty * T + (p / T) % T, i = tz * T + p / (T * T);
        if (i < A.dz && j < A.dy && k < A.dx) out[(size_t(i) * A.dy + j) * A.dx + k] = v;
    }
    if constexpr (VAR) {
        __syncthreads();  // the mean's partial volumes are consumed
#pragma unroll
        for (uint32_t x = 0; x < T; ++x) part[wave][(lz * T + ly) * T + x] = acc_var[x];
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < T * T * T; p += blockDim.x) {
            float v = 0.f;
#pragma unroll
            for (int k = 0; k < kVoxFwdWaves; ++k) v += part[k][p];
            const uint32_t k = tx * T + p % T, j = ty * T + (p / T) % T, i = tz * T + p / (T * T);
            if (i < A.dz && j < A.dy && k < A.dx) out_var[(size_t(i) * A.dy + j) * A.dx + k] = v;
        }
    }
}

// Per-primitive sums of the intersection partials [cum[n], cum[n+1]) in
// emission order (deterministic; no atomics).
__global__ void __launch_bounds__(256)
vox_partials_reduce_kernel(const uint32_t* __restrict__ cum, const float* __restrict__ partials, uint32_t N,
                           uint32_t n_isects, float* __restrict__ v_xyz, float* __restrict__ v_c,
                           float* __restrict__ v_opac, float* __restrict__ v_beta, float* __restrict__ v_unc) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float acc[12] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const uint32_t lo = min(cum[n], n_isects), hi = min(cum[n + 1], n_isects);
    for (uint32_t j = lo; j < hi; ++j) {
        const float4* p = reinterpret_cast<const float4*>(partials + kVoxPartial * size_t(j));
        const float4 a = p[0], b = p[1], c = p[2];
        acc[0] += a.x; acc[1] += a.y; acc[2] += a.z; acc[3] += a.w;
        acc[4] += b.x; acc[5] += b.y; acc[6] += b.z; acc[7] += b.w;
        acc[8] += c.x; acc[9] += c.y; acc[10] += c.z; acc[11] += c.w;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) v_xyz[3 * size_t(n) + k] = acc[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) v_c[6 * size_t(n) + k] = acc[3 + k];
    v_opac[n] = acc[9];
    if (v_beta) v_beta[n] = acc[10];
    if (v_unc) v_unc[n] = acc[11];
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
        if constexpr (std::is_same<S, float>::value) {
            const uint64_t n_tiles = uint64_t(n_tiles_x) * n_tiles_y * n_tiles_z;
            if (tile_size == kMomT && !v_render_var && n_tiles < (1ull << 31)) {  // moments form
                auto km = beta ? vox_bwd_moments_kernel<true, false> : vox_bwd_moments_kernel<false, false>;
                km<<<ceil_div(n_isects, kMomThreads), kMomThreads, 0, stream>>>(
                    A, isect_ids, (const float*)v_render, nullptr, 0.f, (float*)v_means3d, (float*)v_conics,
                    (float*)v_opacities, (float*)v_betas, nullptr, nullptr, nullptr);
                return check_launch("voxelize_to_voxels_bwd_per_gaussian");
            }
        }
        constexpr uint32_t waves = 4;
        auto k = beta ? vox_bwd_isect_kernel<S, true> : vox_bwd_isect_kernel<S, false>;
        k<<<std::min<uint32_t>(ceil_div(n_isects, waves), kIsectBlocks), waves * kWave, 0, stream>>>(
            A, isect_ids, (const S*)v_render, (const S*)v_render_var,
            v_render_var ? (const S*)uncertainties : nullptr, (S*)v_means3d, (S*)v_conics,
            (S*)v_opacities, (S*)v_betas, v_render_var ? (S*)v_uncertainties : nullptr);
        return check_launch("voxelize_to_voxels_bwd_per_gaussian");
    });
}

extern "C" size_t r2s_voxel_partials_bytes(uint32_t n_isects) {
    return size_t(n_isects) * kVoxPartial * sizeof(float);
}

extern "C" int r2s_voxel_frame(const float* means, const float* quats, const float* scales, const double* ranges,
                               const double* centers, uint32_t dim_x, uint32_t dim_y, uint32_t dim_z, double eps,
                               uint32_t N, double* frame, r2s_stream_t stream) {
    if (N == 0) return R2S_OK;
    R2S_REQUIRE(means && quats && scales && ranges && centers && frame, "voxel_frame: null argument");
    vox_frame_kernel<<<ceil_div(N, 256), 256, 0, stream>>>(means, quats, scales, ranges, centers, dim_x, dim_y,
                                                           dim_z, eps, N, frame);
    return check_launch("voxel_frame");
}

extern "C" int r2s_voxelize_fused_vjp(const float* means3d, const double* frame, const float* conics,
                                      const float* opacities,
                                      const float* betas, const float* uncertainties, const int32_t* radii,
                                      const uint32_t* tile_cum, const int64_t* isect_ids, const int32_t* flatten_ids,
                                      uint32_t n_isects, uint32_t N, uint32_t dim_x, uint32_t dim_y, uint32_t dim_z,
                                      uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y, uint32_t n_tiles_z,
                                      const float* v_render, const float* v_render_var, int detach, float* partials,
                                      size_t partials_bytes, float* v_means3d, float* v_conics, float* v_opacities,
                                      float* v_betas, float* v_uncertainties, r2s_stream_t stream) {
    if (int rc = check_vox_tile(tile_size, n_tiles_x, n_tiles_y, n_tiles_z, dim_x, dim_y, dim_z)) return rc;
    R2S_REQUIRE(tile_size == kMomT, "the fused voxel VJP takes %u^3 tiles (got %u)", kMomT, tile_size);
    R2S_REQUIRE(uint64_t(n_tiles_x) * n_tiles_y * n_tiles_z < (1ull << 31), "too many tiles");
    R2S_REQUIRE(partials_bytes >= r2s_voxel_partials_bytes(n_isects), "partials buffer too small: %zu < %zu",
                partials_bytes, r2s_voxel_partials_bytes(n_isects));
    const bool beta = betas != nullptr, var = v_render_var != nullptr;
    R2S_REQUIRE(!beta || v_betas != nullptr, "v_betas is required with betas");
    R2S_REQUIRE(!var || (uncertainties != nullptr && v_uncertainties != nullptr),
                "a variance adjoint needs uncertainties and v_uncertainties");
    R2S_REQUIRE(frame != nullptr, "voxelize_fused_vjp: the double-precision frame is required");
    if (N == 0) return R2S_OK;
    if (n_isects > 0) {
        auto A = make_vox_args<float>(means3d, conics, opacities, betas, uncertainties, nullptr, flatten_ids,
                                      n_isects, N, dim_x, dim_y, dim_z, tile_size, n_tiles_x, n_tiles_y, n_tiles_z);
        A.frame = frame;
        auto km = beta ? (var ? vox_bwd_moments_kernel<true, true, true> : vox_bwd_moments_kernel<true, false, true>)
                       : (var ? vox_bwd_moments_kernel<false, true> : vox_bwd_moments_kernel<false, false>);
        km<<<ceil_div(n_isects, kMomThreads), kMomThreads, 0, stream>>>(
            A, isect_ids, v_render, v_render_var, detach ? 0.f : 2.f, nullptr, nullptr, nullptr, nullptr, radii,
            tile_cum, partials);
        if (int rc = check_launch("voxelize_fused_vjp")) return rc;
    }
    vox_partials_reduce_kernel<<<ceil_div(N, 256), 256, 0, stream>>>(tile_cum, partials, N, n_isects, v_means3d,
                                                                     v_conics, v_opacities, beta ? v_betas : nullptr,
                                                                     var ? v_uncertainties : nullptr);
    return check_launch("voxelize_fused_vjp_reduce");
}

extern "C" int r2s_voxelize_fused_fwd(const float* means3d, const double* frame, const float* conics,
                                      const float* opacities,
                                      const float* betas, const float* uncertainties, const int32_t* tile_offsets,
                                      const int32_t* flatten_ids, uint32_t n_isects, uint32_t N, uint32_t dim_x,
                                      uint32_t dim_y, uint32_t dim_z, uint32_t tile_size, uint32_t n_tiles_x,
                                      uint32_t n_tiles_y, uint32_t n_tiles_z, float* render, float* render_var,
                                      r2s_stream_t stream) {
    if (int rc = check_vox_tile(tile_size, n_tiles_x, n_tiles_y, n_tiles_z, dim_x, dim_y, dim_z)) return rc;
    R2S_REQUIRE(tile_size == kMomT, "the fused voxel forward takes %u^3 tiles (got %u)", kMomT, tile_size);
    const bool var = render_var != nullptr;
    R2S_REQUIRE(!var || uncertainties != nullptr, "render_var requires uncertainties");
    const uint64_t n_tiles = uint64_t(n_tiles_x) * n_tiles_y * n_tiles_z;
    R2S_REQUIRE(n_tiles < (1ull << 31), "too many tiles");
    if (n_tiles == 0) return R2S_OK;
    R2S_REQUIRE(frame != nullptr, "voxelize_fused_fwd: the double-precision frame is required");
    auto A = make_vox_args<float>(means3d, conics, opacities, betas, uncertainties, tile_offsets, flatten_ids,
                                  n_isects, N, dim_x, dim_y, dim_z, tile_size, n_tiles_x, n_tiles_y, n_tiles_z);
    A.frame = frame;
    auto k = betas ? (var ? vox_fwd_fused_kernel<true, true> : vox_fwd_fused_kernel<true, false>)
                   : (var ? vox_fwd_fused_kernel<false, true> : vox_fwd_fused_kernel<false, false>);
    k<<<uint32_t(n_tiles), kVoxFwdWaves * kWave, 0, stream>>>(A, render, render_var);
    return check_launch("voxelize_fused_fwd");
}
